"""Autotuner of the engine's knobs, on the GPU it runs on.

* ``dc_bucket_min``: the divide-and-conquer merge kernel gives one workgroup
  to one tree node and sizes its LDS for the node's power-of-two size class.
  The nodes of a tree level (largest first) are launched in size classes;
  classes below this floor join the smallest launched class (0 = the whole
  level in one launch, every node held to the largest node's LDS and
  occupancy).  The density does not depend on it: only floors that leave it
  bit-identical qualify.
* ``mixed_min_m``: a mixed-precision build reduces the submatrices below this
  dimension in FP64 (no float storage, no refinement sweep) and the others on
  float storage with the FP64 refinement.  Small submatrices gain little from
  the float reduction while the refinement costs them about as much as the
  reduction saves.  Only thresholds whose densities stay within 1e-6 of the
  FP64 build (max per element) qualify.

For every knob the density build of its tuning workloads is timed for every
candidate (best of ``--reps``), the others at their current values; the
candidate with the least summed time, each workload's time relative to its
time at the first candidate, wins.  The table and the choices go to
``ptb_amd/tuning.json``, which ``Engine`` applies at start-up.

Usage: python -m ptb_amd.tune [--knobs dc_bucket_min,mixed_min_m] [--reps 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))

KNOBS = {
    "dc_bucket_min": {"candidates": [0, 64, 128, 256, 512, 1024],
                      "workloads": ["h2o_512", "batch256"], "rule": "identical"},
    "mixed_min_m": {"candidates": [0, 160, 224, 288, 352, 448, 576],
                    "workloads": ["h2o_512", "batch256", "c400h802"], "rule": "accuracy"},
}
P_TOL = 1e-6


def _workload(eng, name):
    import torch
    from . import geometry, surrogate
    mols = geometry.mixed_batch(256, 2) if name == "batch256" else [geometry.config_molecule(name)]
    items = []
    for mol in mols:
        H, S = surrogate.build_hs(mol)
        filt = 1e-4 if mol.name.startswith("h2o") else 1e-5
        dH, dS = torch.from_numpy(H).cuda(), torch.from_numpy(S).cuda()
        torch.cuda.synchronize()
        eng.filter_matrix_sym_dev(dS, mol.ndim, filt)
        eng.filter_matrix_sym_dev(dH, mol.ndim, filt)
        items.append((mol.ndim, mol.nel, dH, dS, torch.zeros_like(dH), mol.n, mol.xyz, mol.aoat))
    torch.cuda.synchronize()
    return items


def _densities(items):
    return [it[4].cpu().numpy().copy() for it in items]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--knobs", default="dc_bucket_min,mixed_min_m")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args(argv)
    import numpy as np
    import torch
    from .engine import Engine
    eng = Engine(0)
    cache = {}
    table, chosen = {}, {}

    def build(items, precision):
        eng.sm_stupid_simple_batch(items, 2, 2, precision=precision, device=True)

    def timed(items, precision):
        build(items, precision)  # warm: tables, arenas
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(args.reps):
            t0 = time.perf_counter()
            build(items, precision)
            torch.cuda.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best

    for knob in args.knobs.split(","):
        spec = KNOBS[knob]
        table[knob] = {}
        for wl in spec["workloads"]:
            if wl not in cache:
                cache[wl] = _workload(eng, wl)
            items = cache[wl]
            ref = None
            if spec["rule"] == "accuracy":
                build(items, "fp64")
                ref = _densities(items)
            table[knob][wl] = {}
            for cand in spec["candidates"]:
                saved = eng.get_tuning(knob)
                eng.set_tuning(knob, cand)
                try:
                    best = timed(items, "mixed")
                    eng.profile(True)
                    build(items, "mixed")
                    merge_ms, levels = eng.profile_query("dc_merge")
                    eng.profile(False)
                    P = _densities(items)
                finally:
                    eng.set_tuning(knob, saved)
                row = {"build_ms": round(best * 1e3, 3)}
                if spec["rule"] == "identical":
                    if ref is None:
                        ref = P
                    row["merge_ms"] = round(merge_ms, 3)
                    row["merge_levels"] = levels
                    row["ok"] = bool(all(np.array_equal(a, b) for a, b in zip(ref, P)))
                else:
                    dp = max(float(np.abs(a - b).max()) for a, b in zip(ref, P))
                    row["dP_max_vs_fp64"] = dp
                    row["ok"] = dp <= P_TOL
                table[knob][wl][str(cand)] = row
                print(f"{knob} {wl} {cand:5d}: {json.dumps(row)}", file=sys.stderr, flush=True)

        def score(c):
            rows = [table[knob][w][str(c)] for w in spec["workloads"]]
            if not all(r["ok"] for r in rows):
                return float("inf")
            first = [table[knob][w][str(spec["candidates"][0])]["build_ms"] for w in spec["workloads"]]
            return sum(r["build_ms"] / f for r, f in zip(rows, first))

        chosen[knob] = min(spec["candidates"], key=score)
        eng.set_tuning(knob, chosen[knob])   # the next knob is tuned with this one in place
        print(f"{knob}: {chosen[knob]}", file=sys.stderr, flush=True)
    out = {"device": torch.cuda.get_device_name(0), "precision": "mixed",
           "knobs": {k: KNOBS[k] for k in table}, "table": table, "chosen": chosen}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(chosen), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner of the engine's size buckets.

The divide-and-conquer merge kernel gives one workgroup to one tree node and
sizes its LDS for the node's power-of-two size class.  The nodes of a tree
level (largest first) are launched in size classes; classes below a floor
join the smallest launched class (``dc_bucket_min``; 0 = the whole level in
one launch, every node held to the largest node's LDS and occupancy).  This
script times the density build of the tuning workloads for every candidate
floor on the current GPU, picks the floor with the least total time, and
writes the table and the choice to ``ptb_amd/tuning.json``, which
``Engine`` applies at start-up.

Usage: python -m ptb_amd.tune [--workloads h2o_512,batch256] [--reps 3]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

CANDIDATES = [0, 64, 128, 256, 512, 1024]
HERE = os.path.dirname(os.path.abspath(__file__))


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="h2o_512,batch256")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--precision", default="mixed")
    ap.add_argument("--out", default=os.path.join(HERE, "tuning.json"))
    args = ap.parse_args(argv)
    import numpy as np
    import torch
    from . import geometry, surrogate
    from .engine import Engine
    eng = Engine(0)
    table = {}
    for wl in args.workloads.split(","):
        mols = geometry.mixed_batch(256, 2) if wl == "batch256" else [geometry.config_molecule(wl)]
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
        table[wl] = {}
        ref = None
        for cand in CANDIDATES:
            eng.set_tuning("dc_bucket_min", cand)
            eng.sm_stupid_simple_batch(items, 2, 2, precision=args.precision, device=True)  # warm
            torch.cuda.synchronize()
            best = float("inf")
            for _ in range(args.reps):
                t0 = time.perf_counter()
                eng.sm_stupid_simple_batch(items, 2, 2, precision=args.precision, device=True)
                torch.cuda.synchronize()
                best = min(best, time.perf_counter() - t0)
            eng.profile(True)
            eng.sm_stupid_simple_batch(items, 2, 2, precision=args.precision, device=True)
            merge_ms, launches = eng.profile_query("dc_merge")
            eng.profile(False)
            P = [it[4].cpu().numpy().copy() for it in items]
            if ref is None:
                ref = P
            same = all(np.array_equal(a, b) for a, b in zip(ref, P))
            table[wl][str(cand)] = {"build_ms": round(best * 1e3, 3), "merge_ms": round(merge_ms, 3),
                                    "merge_levels": launches, "density_bit_identical": same}
            print(f"{wl} dc_bucket_min {cand:5d}: build {best * 1e3:8.2f} ms, merge {merge_ms:7.2f} ms "
                  f"over {launches} tree levels, identical {same}", file=sys.stderr, flush=True)
    # the floor with the least summed build time over the workloads (only
    # floors that leave every density bit-identical qualify)
    def total(c):
        rows = [table[w][str(c)] for w in table]
        return float("inf") if not all(r["density_bit_identical"] for r in rows) else \
            sum(r["build_ms"] for r in rows)
    choice = min(CANDIDATES, key=total)
    out = {"device": torch.cuda.get_device_name(0), "precision": args.precision,
           "candidates": CANDIDATES, "table": table, "chosen": {"dc_bucket_min": choice}}
    with open(args.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out["chosen"]), flush=True)


if __name__ == "__main__":
    main()

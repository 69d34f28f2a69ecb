#!/usr/bin/env python3
"""Density builds per second on one MI355X (512-water box, filter 1e-4).

One step = one density build of the submatrix method (the work of
sm_stupid_simple, source/sm.f90:34-409): filtered packed H and S resident in
HBM -> packed P in HBM, including index sets, gather, per-submatrix solve,
chemical-potential bisection and scatter.  The H/S filter (source/pgtb.f90:
1532-1548) runs once before the timed region and is reported separately.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed|both]
Prints one JSON line on stdout; diagnostics go to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
METRIC = "density builds/s: mixed-prec batched vs FP64 GPU, |ΔE_band|≤1e-6 Eh, 512 H2O"
WORKLOADS = {
    "h2o_512": {"filter": 1e-4, "columns": 2},
    "h2o_64": {"filter": 1e-4, "columns": 2},
    "c400h802": {"filter": 1e-5, "columns": 2},
    "h2o_2048": {"filter": 1e-5, "columns": 2},
}


# bracketed names: whole stages (their GEMMs included) and single kernels
PHASES = ["potrf", "trsm", "trsm_h", "trsm_k", "trsm_e", "trsm_u", "sytrd", "apply_q",
          "apply_q_x", "apply_q_o", "apply_q_r1", "apply_q_r2", "gemm_sytrd", "gemm_potrf",
          "gemm_trsm", "gemm_dc", "gemm_q", "gemm_fe", "gemm_refine", "gemm_other", "refine_all",
          "inv_lower", "gemm_linv", "gemm_kform", "gemm_u"]
KERNELS = ["pattern", "group_union", "group_map", "gather", "potrf_panel", "diag_inv",
           "transpose", "sytrd_panel", "sytrd_panel_f32", "zero_upper", "dc_leaf", "dc_merge", "gemm", "larft",
           "fe_rows", "fe_finish", "bisect", "occ_compact", "dmat_scatter", "to_f32", "promote",
           "refine"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sytrd_bytes(ms, elem=8):
    """Algorithmic HBM bytes of the Householder tridiagonalisation of one build
    (all panel launches): the symmetric trailing matrix (one triangle) read
    once per reflector, elem * sum_k (m-k)(m-k+1)/2 per submatrix (elem = 8
    for FP64 storage, 4 for the float storage of the mixed-precision solve)."""
    tot = 0.0
    for m in ms:
        n = np.arange(m - 1, 1, -1, dtype=np.float64)  # trailing sizes m-1 .. 2
        tot += float(np.sum(n * (n + 1) / 2.0)) * elem
    return tot


def roofline_model(name, ms, ndim, mixed=False, thr=0):
    """(bound, algorithmic units of one build summed over the kernel's launches,
    unit); divided by the kernel's bracketed time of one build this is the
    per-launch figure averaged over its launches.  A mixed build reduces the
    submatrices below `thr` (mixed_min_m) with the FP64 panel kernel
    (`sytrd_panel`) and the others with the float one (`sytrd_panel_f32`)."""
    m = np.asarray(ms, dtype=np.float64)
    npk = ndim * (ndim + 1) / 2.0
    if name == "sytrd_panel_f32":
        return "hbm", sytrd_bytes([x for x in ms if x >= thr], 4), "GB/s"
    if name == "sytrd_panel":
        sel = [x for x in ms if not mixed or x < thr]
        return ("hbm", sytrd_bytes(sel, 8), "GB/s") if sel else None
    if name == "gather":   # dense Hsm and Ssm written + packed reads counted once
        return "hbm", float(np.sum(16.0 * m * m)) + 16.0 * npk, "GB/s"
    if name == "pattern":
        return "hbm", 16.0 * npk, "GB/s"
    if name == "filter":
        return "hbm", 8.0 * npk, "GB/s"
    return None


def traffic_per_launch(kernel, workload, mixed=False):
    """Scaled HBM bytes per dispatch of `kernel` (its float-storage instance for
    the mixed-precision solve) from the committed rocprofv3 --pmc summary of
    this workload (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's
    reading for a plain copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_hbm_traffic.md")
    key = "bytes_per_launch_float=" if mixed else "bytes_per_launch="
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key):
                    return float(line.split("=", 1)[1])
    except OSError:
        return None
    return None


BF16_PEAK_TFS = 2500.0   # dense BF16 matrix-core peak; split-BF16x3 runs 3 products per step


def cpu_info():
    """CPU model, the cores this process may use, and the BLAS numpy calls."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    blas = "unknown"
    try:
        from threadpoolctl import threadpool_info
        for lib in threadpool_info():
            if lib.get("user_api") == "blas":
                blas = f"{lib.get('internal_api')} {lib.get('version')} ({os.path.basename(lib.get('filepath', ''))})"
                break
    except Exception:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return {"cpu_model": model, "cores_available": avail, "blas": blas}


def gemm_site_flops(site, ms, cs, nos, mixed, thr=0, lowp_max=0):
    """Algorithmic flops of one build's products at a GEMM site, from the
    submatrix dimensions m, owned columns c and occupied states n_o, as a pair
    (FP64 flops, split-BF16x3 flops).  In the mixed build the submatrices at
    or above `thr` (mixed_min_m) are refined; of those, the ones up to
    `lowp_max` rows (lowp_merge_max_m) take split-BF16x3 merge products and a
    second Loewdin step.  The refinement's correction products (Z^T R and
    Z E) are split-BF16x3, its residual and Loewdin products FP64."""
    m = np.asarray(ms, dtype=np.float64)
    c = np.asarray(cs, dtype=np.float64)
    no = np.asarray(nos, dtype=np.float64)
    refined = (m >= thr) if mixed else np.zeros(m.shape, dtype=bool)
    lowp = refined & (m <= lowp_max)
    if site == "gemm_potrf":
        return float(np.sum(m ** 3) / 3.0), 0.0
    if site == "gemm_linv":
        return float(np.sum(m ** 3) / 3.0), 0.0
    if site == "gemm_kform":
        return float(np.sum(m ** 3)), 0.0
    if site == "gemm_sytrd":
        return float(np.sum(2.0 * m ** 3 / 3.0)), 0.0
    if site == "gemm_dc":
        f = 2.0 * m ** 3 / 3.0
        return float(np.sum(f[~lowp])), float(np.sum(f[lowp]))
    if site == "gemm_u":
        return float(np.sum(m * m * no)), 0.0
    if site == "gemm_fe":
        return float(np.sum(2.0 * m * m * 2.0 * c)) * (2.0 if mixed else 1.0), 0.0
    if site == "gemm_q":
        f = np.sum(4.0 * m * m * (2.0 * c + no))
        return float(f + np.sum((8.0 * m * m * no)[refined])), 0.0
    if site == "gemm_refine":
        steps = 2.0 if lowp.any() else 1.0
        fp = (2.0 * m * m * no + steps * 2 * 2.0 * m * no * no)[refined]
        bf = (2 * 2.0 * m * m * no)[refined]
        return float(np.sum(fp)), float(np.sum(bf))
    return None


def cpu_baseline_full_h2o64(budget_note=""):
    """One complete oracle build (numpy/scipy restatement of sm_stupid_simple)
    of the 64-water box at filter 1e-4, atoms_columns: the reference's own
    CPU-runnable case."""
    from ptb_amd import geometry, surrogate
    from oracle import ptb_oracle as O
    mol = geometry.config_molecule("h2o_64")
    H, S = surrogate.build_hs(mol)
    O.filter_matrix_sym(S, mol.ndim, 1e-4)
    O.filter_matrix_sym(H, mol.ndim, 1e-4)
    t0 = time.perf_counter()
    O.sm_stupid_simple(mol.ndim, mol.nel, H, S, mol.n, mol.xyz, mol.aoat, 2, 2)
    t = time.perf_counter() - t0
    return {"value": 1.0 / t, "unit": "builds/s", "seconds_per_build": round(t, 2),
            "workload": "h2o_64 (64 H2O, ndim 1472, 192 submatrices), filter 1e-4, atoms_columns",
            "kind": "port", "sample": "one complete build, mu bisection, scatter and pack included"}


def cpu_baseline(mol, H, S, cols, budget_s=20.0):
    """The oracle (numpy/scipy restatement of source/sm.f90) timed on a bounded
    sample of the same workload: every submatrix's gather + dsygvd + fe +
    owned-column dmat for an evenly spaced subset, scaled to all submatrices,
    plus the blow-up and index-set setup measured once."""
    from threadpoolctl import threadpool_limits
    from oracle import ptb_oracle as O
    cores = max(1, min(16, os.cpu_count() or 1))
    with threadpool_limits(limits=cores):
        t0 = time.perf_counter()
        n = mol.ndim
        Hd = O.blowsym(H, n)
        Sd = O.blowsym(S, n)
        pat = O.pattern_dense(Hd, Sd)
        groups = O.column_groups(cols, n, mol.aoat, mol.n, mol.xyz, pat)
        t_setup = time.perf_counter() - t0
        nsm = len(groups)
        order = np.linspace(0, nsm - 1, num=nsm).astype(int)
        stride = max(1, nsm // 64)
        sample = list(order[::stride])
        done, t_sm = 0, 0.0
        for s in sample:
            g = groups[s]
            t1 = time.perf_counter()
            rows = np.nonzero(pat[:, g].any(axis=1))[0]
            imap = np.full(n, -1)
            imap[rows] = np.arange(len(rows))
            Hs = Hd[np.ix_(rows, rows)]
            Ss = Sd[np.ix_(rows, rows)]
            e, U, _ = O.sm_eigendecomp(Hs, Ss)
            O.sm_fe(Ss, U, imap[g])
            occ = O.occupations(e, -0.5)
            O.dmat(occ, U)
            t_sm += time.perf_counter() - t1
            done += 1
            if t_setup + t_sm > budget_s:
                break
    per_sm = t_sm / done
    est = t_setup + per_sm * nsm
    info = cpu_info()
    return {"value": 1.0 / est, "unit": "builds/s", "cores": cores, "kind": "port",
            "sample": f"extrapolated: oracle (scipy dsygvd, {cores} threads) timed on {done} of "
                      f"{nsm} submatrices ({t_sm:.1f} s) + dense setup ({t_setup:.1f} s), scaled "
                      f"to all submatrices: {est:.1f} s per build (no bisection, scatter or pack)",
            **info}


def run_batch(args):
    """The batch of 256 independent molecules (half water clusters of 4-40
    H2O at filter 1e-4, half CnH2n+2 with n in 3-40 at filter 1e-5, sizes from
    seed 2): K timed batched builds per precision (molecules/s), and every
    molecule built alone once (p50/p99 of the single-molecule time)."""
    import torch
    from ptb_amd import geometry, surrogate
    from ptb_amd.engine import Engine
    mols = geometry.mixed_batch(256, 2)
    eng = Engine(0)
    stream = torch.cuda.ExternalStream(eng.stream_handle())
    items, host = [], []
    for mol in mols:
        H, S = surrogate.build_hs(mol)
        filt = 1e-4 if mol.name.startswith("h2o") else 1e-5
        eng.filter_matrix_sym(S, mol.ndim, filt)
        eng.filter_matrix_sym(H, mol.ndim, filt)
        host.append((mol, H))
        items.append((mol.ndim, mol.nel, torch.from_numpy(H).cuda(), torch.from_numpy(S).cuda(),
                      torch.zeros(H.size, dtype=torch.float64, device="cuda"), mol.n, mol.xyz,
                      mol.aoat))
    torch.cuda.synchronize()
    log(f"batch: {len(items)} molecules, ndim {min(m.ndim for m in mols)}..{max(m.ndim for m in mols)}")
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    res, Ps = {}, {}
    for prec in precisions:
        for _ in range(args.warmup):
            eng.sm_stupid_simple_batch(items, 2, 2, precision=prec, device=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            reps = eng.sm_stupid_simple_batch(items, 2, 2, precision=prec, device=True)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / args.steps
        Ps[prec] = [it[4].cpu().numpy().copy() for it in items]
        single = []
        for it in items:
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            eng.sm_stupid_simple_dev(*it, 2, 2, precision=prec)
            torch.cuda.synchronize()
            single.append((time.perf_counter() - t1) * 1000.0)
        # the same batch with the eigen phase replayed from a captured hipGraph
        # (one build to see the shapes, one to capture, then K timed replays)
        eng.set_tuning("graph", 1)
        for _ in range(2):
            eng.sm_stupid_simple_batch(items, 2, 2, precision=prec, device=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            eng.sm_stupid_simple_batch(items, 2, 2, precision=prec, device=True)
        torch.cuda.synchronize()
        gms = (time.perf_counter() - t0) * 1000.0 / args.steps
        same = all(np.array_equal(a, it[4].cpu().numpy()) for a, it in zip(Ps[prec], items))
        held = eng.get_tuning("graphs")
        eng.set_tuning("graph", 0)
        res[prec] = {"ms_per_batch": round(ms, 3), "molecules_per_s": round(len(items) * 1000.0 / ms, 2),
                     "graph": {"ms_per_batch": round(gms, 3),
                               "molecules_per_s": round(len(items) * 1000.0 / gms, 2),
                               "graphs_held": held, "bit_identical_to_eager": same},
                     "single_p50_ms": round(float(np.percentile(single, 50)), 3),
                     "single_p99_ms": round(float(np.percentile(single, 99)), 3),
                     "single_sum_ms": round(float(np.sum(single)), 1),
                     "nsm": int(sum(r["nsm"] for r in reps)),
                     "dc_bucket_min": eng.get_tuning("dc_bucket_min")}
        log(f"[{prec}] {ms:.1f} ms per batch of {len(items)}, single-molecule p50 "
            f"{res[prec]['single_p50_ms']} ms p99 {res[prec]['single_p99_ms']} ms "
            f"(sum {res[prec]['single_sum_ms']} ms)")
    main_prec = "mixed" if "mixed" in res else "fp64"
    r = res[main_prec]
    out = {"metric": "molecule density builds/s, batch of 256 independent molecules",
           "value": r["molecules_per_s"], "unit": "molecules/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_batch"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64" if main_prec == "fp64" else "f32+f64",
           "data": "synthetic (deterministic geometries, seed 2; Gaussian/extended-Hueckel surrogate H,S)",
           "config": {"workload": "batch256 (128 water clusters of 4-40 H2O at filter 1e-4, 128 "
                                  "CnH2n+2 n=3-40 at filter 1e-5), atoms_columns, submatrix_sygv",
                      "precision": main_prec, "mixed_min_m": eng.get_tuning("mixed_min_m"),
                      "dc_bucket_min": eng.get_tuning("dc_bucket_min"),
                      "gemm_tile": eng.get_tuning("gemm_tile"),
                      "dc_global_min": eng.get_tuning("dc_global_min"),
                      "lowp_merge_max_m": eng.get_tuning("lowp_merge_max_m")},
           "per_precision": res}
    if "fp64" in Ps and "mixed" in Ps:
        dp, de = 0.0, 0.0
        for (mol, H), a, b in zip(host, Ps["fp64"], Ps["mixed"]):
            dp = max(dp, float(np.abs(a - b).max()))
            de = max(de, abs(eng.band_structure(mol.ndim, a, H) - eng.band_structure(mol.ndim, b, H)))
        out["mixed_vs_fp64"] = {"speedup": round(res["fp64"]["ms_per_batch"] / res["mixed"]["ms_per_batch"], 3),
                                "dP_max": dp, "dE_band_max": de}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="h2o_512", choices=sorted(WORKLOADS) + ["batch256"])
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true",
                    help="skip the dense -check solve (solve2) of the 512-water box")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="run one plain device copy of this many MiB first (calibrates HBM "
                         "byte counters under rocprofv3 --pmc)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the density build "
              "is not sharded across GPUs in this engine", file=sys.stderr)
        sys.exit(2)

    if args.config == "batch256":
        run_batch(args)
        return

    import torch
    from ptb_amd import geometry, surrogate
    from ptb_amd.engine import Engine

    if args.copy_mib > 0:
        src = torch.ones(args.copy_mib * 131072, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.copy_mib} MiB read + {args.copy_mib} MiB written")
        del src, dst
    wl = WORKLOADS[args.config]
    mol = geometry.config_molecule(args.config)
    H, S = surrogate.build_hs(mol)
    n = mol.ndim
    log(f"workload {args.config}: atoms {mol.n} ndim {n} nel {mol.nel}")
    eng = Engine(0)
    stream = torch.cuda.ExternalStream(eng.stream_handle())
    dH = torch.from_numpy(H).cuda()
    dS = torch.from_numpy(S).cuda()
    torch.cuda.synchronize()
    eng.profile(True)
    cS = eng.filter_matrix_sym_dev(dS, n, wl["filter"])
    cH = eng.filter_matrix_sym_dev(dH, n, wl["filter"])
    filt_ms, filt_n = eng.profile_query("filter")          # the call, count readback included
    fk_ms, fk_n = eng.profile_query("filter_kernel")      # the packed pass alone
    if fk_n == 0:
        fk_ms, fk_n = filt_ms, filt_n
    eng.profile(False)
    log(f"filter: S positive {cS}, H positive {cH}, {filt_ms / filt_n:.3f} ms per matrix")
    dP = torch.zeros_like(dH)
    precisions = ["fp64", "mixed"] if args.precision == "both" else [args.precision]
    results = {}
    for prec in precisions:
        for _ in range(args.warmup):
            eng.sm_stupid_simple_dev(n, mol.nel, dH, dS, dP, mol.n, mol.xyz, mol.aoat,
                                     wl["columns"], 2, precision=prec)
        torch.cuda.synchronize()
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ev0.record(stream)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            rep = eng.sm_stupid_simple_dev(n, mol.nel, dH, dS, dP, mol.n, mol.xyz, mol.aoat,
                                           wl["columns"], 2, precision=prec)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        gpu_ms = ev0.elapsed_time(ev1)
        # one more build with the per-kernel HIP-event brackets on (outside the
        # timed region: the brackets cost host time of their own)
        eng.profile(True)
        eng.sm_stupid_simple_dev(n, mol.nel, dH, dS, dP, mol.n, mol.xyz, mol.aoat,
                                 wl["columns"], 2, precision=prec)
        kern = {}
        for name in PHASES + KERNELS:
            ms, cnt = eng.profile_query(name)
            if cnt:
                kern[name] = (ms, cnt)
        eng.profile(False)
        ms_step = wall * 1000.0 / args.steps
        m, _, _ = eng.sm_spectra(rep["nsm"], rep["sum_m"])
        results[prec] = {"ms_per_step": ms_step, "gpu_ms_per_step": gpu_ms / args.steps,
                         "rep": rep, "kernels": kern, "m": m,
                         "P": dP.clone() if n > 20000 else dP.cpu().numpy().copy()}
        log(f"[{prec}] {ms_step:.1f} ms/build (stream {gpu_ms / args.steps:.1f} ms), mu {rep['mu']}, "
            f"nsm {rep['nsm']}, max m {rep['max_m']}")
        for k, (ms, cnt) in sorted(kern.items(), key=lambda kv: -kv[1][0]):
            log(f"    {k:14s} {ms:9.3f} ms/build  ({cnt} launch groups)")

    main_prec = "mixed" if "mixed" in results else "fp64"
    r = results[main_prec]
    value = 1000.0 / r["ms_per_step"]
    thr = eng.get_tuning("mixed_min_m")
    # roofline of the dominant kernel that has a model
    roof = None
    for name, (ms, cnt) in sorted(r["kernels"].items(), key=lambda kv: -kv[1][0]):
        if name not in KERNELS:
            continue
        model = roofline_model(name, r["m"], n, mixed=main_prec == "mixed", thr=thr)
        if model is None:
            continue
        bound, units, unit = model
        # units and ms both cover all of the kernel's launches in one build, so
        # units / ms is the per-launch figure averaged over its launches
        per_launch_ms = ms / max(cnt, 1)
        if bound == "hbm":
            achieved = units / (ms * 1e-3) / 1e9
            peak = HBM_PEAK_GBS
        else:
            achieved = units / (ms * 1e-3) / 1e12
            peak = FP64_PEAK_TFS
        kname = "k_sytrd_panel" if name.startswith("sytrd_panel") else "k_" + name
        f32 = name == "sytrd_panel_f32"
        traffic = traffic_per_launch(kname, args.config, mixed=f32)
        roof = {"kernel": kname + ("<float>" if f32 else "<double>" if name == "sytrd_panel" else ""),
                "bound": bound, "achieved": round(achieved, 2), "peak": peak,
                "unit": unit, "frac": round(achieved / peak, 4),
                "traffic": None if traffic is None else round(traffic, 0),
                "algorithmic_bytes_per_launch": round(units / max(cnt, 1), 0),
                "ms_per_launch": round(per_launch_ms, 3), "share_of_build": round(ms / r["ms_per_step"], 3)}
        break
    # per-site rooflines of the main build: every GEMM site against the peak of
    # its mix of arithmetic (the FP64 matrix cores for FP64 products, a third
    # of the BF16 peak for split-BF16x3 ones), the HBM-bound kernels in GB/s
    starts = np.r_[0, np.flatnonzero(np.diff(mol.aoat)) + 1, n]
    cs = np.diff(starts) if wl["columns"] == 2 else None
    sites = []
    if cs is not None and len(cs) == len(r["m"]):
        _, eigs, _ = eng.sm_spectra(r["rep"]["nsm"], r["rep"]["sum_m"])
        nos = np.array([int(np.sum(e < r["rep"]["mu"])) for e in eigs])
        bf_peak = BF16_PEAK_TFS / 3.0
        for site, (ms, cnt) in sorted(r["kernels"].items(), key=lambda kv: -kv[1][0]):
            fl = gemm_site_flops(site, r["m"], cs, nos, main_prec == "mixed", thr,
                                 eng.get_tuning("lowp_merge_max_m"))
            if fl is None or ms <= 0:
                continue
            fp, bf = fl
            tf = (fp + bf) / (ms * 1e-3) / 1e12
            # the site's peak for its mix of arithmetic: its flops over the
            # least time the two kinds of matrix cores need for them
            peak = (fp + bf) / (fp / FP64_PEAK_TFS + bf / bf_peak)
            if bf == 0:
                arith = "fp64"
            elif fp == 0:
                arith = "bf16x3 (3 bf16 products)"
            else:
                arith = f"fp64 + bf16x3 ({100.0 * bf / (fp + bf):.0f} % of the flops)"
            sites.append({"site": site, "bound": "mfma", "arith": arith,
                          "algorithmic_tflop": round((fp + bf) / 1e12, 4), "ms": round(ms, 3),
                          "achieved": round(tf, 2), "peak": round(peak, 1),
                          "unit": "TFLOP/s", "frac": round(tf / peak, 4)})
        for name in ("sytrd_panel_f32", "sytrd_panel", "gather", "pattern"):
            if name in r["kernels"]:
                ms, cnt = r["kernels"][name]
                model = roofline_model(name, r["m"], n, mixed=main_prec == "mixed", thr=thr)
                if model is None:
                    continue
                bound, units, unit = model
                gbs = units / (ms * 1e-3) / 1e9
                sites.append({"site": name, "bound": bound, "algorithmic_gb": round(units / 1e9, 4),
                              "ms": round(ms, 3), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                              "unit": unit, "frac": round(gbs / HBM_PEAK_GBS, 4)})
        if "dmat_scatter" in r["kernels"]:
            ms, _ = r["kernels"]["dmat_scatter"]
            mm = np.asarray(r["m"], dtype=np.float64)
            # owned columns of U_o read (m x n_o per submatrix... the c rows of
            # it the columns need) and the packed P written once
            units = float(np.sum(8.0 * mm * nos)) + 8.0 * n * (n + 1) / 2.0
            gbs = units / (ms * 1e-3) / 1e9
            sites.append({"site": "dmat_scatter", "bound": "hbm", "algorithmic_gb": round(units / 1e9, 4),
                          "ms": round(ms, 3), "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                          "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)})
        gbs = 16.0 * n * (n + 1) / 2.0 / (fk_ms / max(fk_n, 1) * 1e-3) / 1e9
        sites.append({"site": "filter", "bound": "hbm", "algorithmic_gb": round(16.0 * n * (n + 1) / 2.0 / 1e9, 4),
                      "ms": round(fk_ms / max(fk_n, 1), 3), "achieved": round(gbs, 1),
                      "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)})
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_prec == "fp64" else "f32+f64",
        "data": "synthetic (deterministic geometry; Gaussian/extended-Hueckel surrogate H,S)",
        "config": {"workload": f"{args.config} ({mol.name}, {mol.n} atoms, ndim {n}), filter {wl['filter']}, "
                               f"atoms_columns, submatrix_sygv",
                   "nsm": int(r["rep"]["nsm"]), "max_m": int(r["rep"]["max_m"]),
                   "sum_m3": float(r["rep"]["sum_m3"]), "precision": main_prec,
                   "mixed_min_m": eng.get_tuning("mixed_min_m"),
                   "dc_bucket_min": eng.get_tuning("dc_bucket_min"),
                   "gemm_tile": eng.get_tuning("gemm_tile"),
                   "dc_global_min": eng.get_tuning("dc_global_min"),
                   "lowp_merge_max_m": eng.get_tuning("lowp_merge_max_m")},
        "fp64_builds_per_s": round(1000.0 / results["fp64"]["ms_per_step"], 4) if "fp64" in results else None,
        "filter_ms_per_matrix": round(filt_ms / max(filt_n, 1), 4),
        "filter_kernel_ms_per_matrix": round(fk_ms / max(fk_n, 1), 4),
        "roofline": roof,
        "roofline_sites": sites,
    }
    big = n > 20000   # the 2048-water box: no host copies of its 9 GB matrices
    Hf = dH.cpu().numpy() if not big else None
    out["filter_plus_build_ms"] = round(2.0 * filt_ms / max(filt_n, 1) + r["ms_per_step"], 3)
    if not big:
        # the drop-in on host arrays (PCIe transfers of packed H, S in and P
        # out included); not `value`
        Sf = dS.cpu().numpy()
        Ph = np.empty_like(Hf)
        t0 = time.perf_counter()
        eng.sm_stupid_simple(n, mol.nel, Hf, Sf, Ph, mol.n, mol.xyz, mol.aoat, wl["columns"], 2,
                             precision=main_prec)
        out["host_arrays_ms_per_build"] = round((time.perf_counter() - t0) * 1000.0, 3)
    if "fp64" in results and "mixed" in results:
        P64, Pm = results["fp64"]["P"], results["mixed"]["P"]
        if big:
            e64 = eng.band_structure_dev(n, P64, dH)
            em = eng.band_structure_dev(n, Pm, dH)
            dpmax = float((P64 - Pm).abs().max().item())
        else:
            e64 = eng.band_structure(n, P64, Hf)
            em = eng.band_structure(n, Pm, Hf)
            dpmax = float(np.abs(P64 - Pm).max())
        out["mixed_vs_fp64"] = {
            "speedup": round(results["fp64"]["ms_per_step"] / results["mixed"]["ms_per_step"], 3),
            "dP_max": dpmax, "dE_band": float(abs(em - e64))}
        # the precision ladder: the mixed build with its divide-and-conquer
        # merge products (submatrices up to 1024 rows) in split-BF16x3 and in
        # plain BF16 (8-bit operands), against FP64 (the shipped mixed build
        # takes FP64 merge products and BF16x3 correction products)
        if not big:
            saved = eng.get_tuning("lowp_merge_max_m")
            eng.set_tuning("lowp_merge_max_m", 1024)
            ladder = {}
            try:
                for label, prec in (("bf16x3", "mixed"), ("bf16", "mixed_bf16")):
                    eng.sm_stupid_simple_dev(n, mol.nel, dH, dS, dP, mol.n, mol.xyz, mol.aoat,
                                             wl["columns"], 2, precision=prec)
                    Pb = dP.cpu().numpy()
                    ladder[label] = {"dP_max": float(np.abs(P64 - Pb).max()),
                                     "dE_band": float(abs(eng.band_structure(n, Pb, Hf) - e64))}
            finally:
                eng.set_tuning("lowp_merge_max_m", saved)
            out["precision_errors"] = {"merge_products": "up to 1024 rows", **ladder}
        # the metric is quoted at |dE_band| <= 1e-6 Eh (and max |dP| <= 1e-6):
        # a mixed build outside that bound posts no throughput
        acc = out["mixed_vs_fp64"]
        if main_prec == "mixed" and (acc["dP_max"] > 1e-6 or acc["dE_band"] > 1e-6):
            out["value"] = None
            out["accuracy_gate"] = "failed: mixed build outside |dE_band|, max|dP| <= 1e-6"
            log(f"accuracy gate failed: dP_max {acc['dP_max']:.3e}, dE_band {acc['dE_band']:.3e}")
    if args.config == "h2o_512" and not args.no_check:
        # the -check reference density (solve2, source/pgtb.f90:1066-1170) of
        # the whole box: one dense generalized eigenproblem of ndim 11776,
        # from host arrays (U returned, as the reference keeps it).  On the
        # unfiltered pencil: the filtered overlap of this box is indefinite.
        try:
            Hc, Sc = surrogate.build_hs(mol)
            t0 = time.perf_counter()
            Pc, ec, fc, Uc, fail = eng.solve2(1, n, mol.nel, 0, 300.0, Hc, Sc)
            ms = (time.perf_counter() - t0) * 1000.0
            out["check_solve2"] = {"ms": round(ms, 1), "ndim": n, "fail": bool(fail),
                                   "tr_PS": round(eng.get_nel(n, Pc, Sc), 9), "nel": mol.nel,
                                   "pencil": "unfiltered H, S of the same box"}
            log(f"solve2 (-check) ndim {n}: {ms:.1f} ms, fail {fail}")
            del Hc, Sc, Pc, Uc
        except Exception as exc:  # reported, never fatal
            out["check_solve2"] = {"error": str(exc)}
    if not args.no_cpu and not big:
        try:
            out["cpu_baseline"] = cpu_baseline(mol, Hf, Sf, wl["columns"], args.cpu_budget)
        except Exception as exc:  # reported, never fatal
            out["cpu_baseline"] = {"error": str(exc)}
        try:
            from threadpoolctl import threadpool_limits
            with threadpool_limits(limits=max(1, min(16, os.cpu_count() or 1))):
                out["cpu_baseline_h2o64_full"] = cpu_baseline_full_h2o64()
        except Exception as exc:
            out["cpu_baseline_h2o64_full"] = {"error": str(exc)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

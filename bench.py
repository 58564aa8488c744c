#!/usr/bin/env python3
"""Benchmark: end-to-end clump finding on one MI355X.

Workload: the 16,777,216-particle (256^3) snapshot of the plain drop-in
configuration — box [-0.1, 0.1]^3, 90% uniform background, 10% in 200
Plummer clumps (masses log-uniform over two decades, scale radii
log-uniform in [5e-4, 4e-3]), float32 positions/velocities, K = 32 (H auto-
raises to 64), seed 0, synthetic.

One step = one pass of the path over the resident snapshot: restore the
N input particles (device copy), MakeGhostParticles, BuildTree,
FindPlanetesimals (densities, HOP, merging, unbinding, rebinding, filters,
angular momentum) to a catalog in host memory.  Inputs are in HBM when the
timed region starts (value); the same steps starting from host memory, with
the upload inside each step, are reported beside it (with_upload).

Output: one JSON line (metric, value = particles/s, roofline of the dominant
kernel — the density KNN launch — with achieved algorithmic GB/s measured
by HIP events on the launch's own stream, and the CPU baseline: the oracle's
OpenMP restatement on the same snapshot, and on the 262,144-particle
configuration, median of three).
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

PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
# useful flops per pair term: 3 sub, r^2 (1 mul + 2 FMA = 5), rsqrt 1, x m 1, accumulate 1
PAIR_FLOPS = 11
# algorithmic bytes per tree particle for the density pass: binary64 position
# (24) + mass (8) read, density (8) written
DENSITY_BYTES_PER_PARTICLE = 40
TRAFFIC_SUMMARY = "knn_select_density_c2_16m_pmc.md"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(launch_units: int):
    """HBM bytes per density launch from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE runs, each divided by the factor the
    same counter reads for a device copy of known size), or None when the
    summary is absent or was taken at another launch size."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_SUMMARY)
    if not os.path.exists(path):
        return None
    vals = {}
    with open(path) as f:
        for line in f:
            if ":" in line and line.startswith("- "):
                k, v = line[2:].split(":", 1)
                vals[k.strip()] = v.strip().split()[0]
    try:
        if int(vals["launch_units"]) != launch_units:
            return None
        return float(vals["hbm_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def boundness_bench(ctx, args, params) -> dict:
    """The second half of the metric: pairwise boundness clumps/s, FP32
    compensated + batched against FP64 per clump, on the batching workload
    (5,000 Plummer clumps, 32..2e6 members with dN/dlnN ~ N^-0.6, 0-30%
    hot-shell interlopers, seed 0); and the precision form on the catalog
    of the last clump-finding step (energy error against FP64 per clump)."""
    import plan_amd
    from plan_amd import synth

    out = {}
    # precision form on the clumps just found in the 16.8M snapshot, 5% of
    # them rescaled to E_kin/|E_pot| in [0.999, 1.001]; errors per clump
    # against the binary64 per-clump form (itself checked against the FP64
    # CPU oracle by the tests)
    ctx.restore_input()
    ctx.make_ghosts()
    ctx.build_tree(True)
    cat = ctx.find_planetesimals()
    parts = ctx.tree_particles()
    off = np.zeros(len(cat["clumps"]) + 1, np.uint32)
    mem = []
    for k, c in enumerate(cat["clumps"]):
        m = cat["members"][c["member_offset"]:c["member_offset"] + c["num_members"]]
        mem.append(m)
        off[k + 1] = off[k] + len(m)
    mem = np.concatenate(mem)
    G = synth.grav_constant(params)
    batch = synth.ClumpBatch(offsets=off, pos=parts["pos"][mem].astype(np.float32),
                             vel=parts["vel"][mem].astype(np.float32), mass=parts["mass"][mem],
                             n_interlopers=np.zeros(len(cat["clumps"]), np.int64), G=G)
    rng = np.random.default_rng(0)
    nseed = max(1, int(round(0.05 * len(cat["clumps"]))))
    seeded = rng.choice(len(cat["clumps"]), nseed, replace=False)
    for c in seeded:
        synth.rescale_to_virial(batch, int(c), float(rng.uniform(0.999, 1.001)))
    ctx.boundness_load(batch)
    s64 = ctx.boundness_run(ctx.FP64)
    _, r64 = ctx.boundness_result()
    s32 = ctx.boundness_run(ctx.FP32)
    _, r32 = ctx.boundness_result()
    ok = (r32["n_bound"] == r64["n_bound"]) & (r64["n_bound"] > 1)
    err = np.maximum(np.abs(r32["E_pot"] - r64["E_pot"]) / np.abs(r64["E_pot"]),
                     np.abs(r32["E_kin"] - r64["E_kin"]) / np.abs(r64["E_kin"]))
    out["precision_form_on_catalog"] = {
        "clumps": len(cat["clumps"]), "members": int(off[-1]), "seeded_near_virial": [int(c) for c in seeded],
        "same_bound_count": int(ok.sum()),
        "per_clump_rel_err": [float(e) if k else None for e, k in zip(err, ok)],
        "rel_err_p50": float(np.median(err[ok])) if ok.any() else None,
        "rel_err_max": float(err[ok].max()) if ok.any() else None,
        "fp64_ms": s64["ms"], "fp32_ms": s32["ms"], "refit_clumps": s32["refit_clumps"]}
    log(f"[bench] boundness on catalog: {out['precision_form_on_catalog']}")
    # batching workload
    t0 = time.perf_counter()
    batch = synth.boundness_batch(n_clumps=args.boundness_clumps, seed=0)
    log(f"[bench] boundness batch: {batch.num_clumps} clumps, {int(batch.offsets[-1])} members, "
        f"generated in {time.perf_counter() - t0:.1f}s")
    ctx.boundness_load(batch)
    ctx.boundness_run(ctx.FP32)  # warm-up: autotuner and graph
    st32 = ctx.boundness_run(ctx.FP32)
    _, b32 = ctx.boundness_result()
    log(f"[bench] fp32 batched: {st32}")
    st64 = ctx.boundness_run(ctx.FP64)
    _, b64 = ctx.boundness_result()
    log(f"[bench] fp64 per clump: {st64}")
    ok = (b32["n_bound"] == b64["n_bound"]) & (b64["n_bound"] > 1)
    rel = np.abs(b32["E_pot"][ok] - b64["E_pot"][ok]) / np.abs(b64["E_pot"][ok])
    nc = batch.num_clumps
    tflops = st32["pair_evaluations"] * PAIR_FLOPS / (st32["ms"] * 1e-3) / 1e12
    out.update({
        "metric": "boundness clumps/s FP32 vs FP64",
        "workload": f"{nc} Plummer clumps, 32..2e6 members (dN/dlnN ~ N^-0.6), 0-30% interlopers, seed 0",
        "members": int(batch.offsets[-1]),
        "fp32_batched_clumps_per_s": nc / (st32["ms"] * 1e-3),
        "fp64_per_clump_clumps_per_s": nc / (st64["ms"] * 1e-3),
        "speedup": st64["ms"] / st32["ms"],
        "fp32_ms": st32["ms"], "fp64_ms": st64["ms"],
        "same_bound_count": int(ok.sum()),
        "max_rel_err_E_pot_vs_fp64": float(rel.max()) if ok.any() else None,
        "pair_evaluations": st32["pair_evaluations"],
        "tile_shape": st32["tile_shape"], "autotune_ms": st32["tune_ms"], "graph_replays": st32["graph_replays"],
        "refit_clumps": st32["refit_clumps"],
        "roofline": {"kernel": "k_bnd_pairs<float> (pairwise potential, Kahan)", "bound": "valu",
                     "achieved": tflops, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / PEAK_FP32_TFLOPS, "flops_per_pair": PAIR_FLOPS,
                     "note": "stage time incl. lists, totals and energies; pair flops only"},
    })
    return out


def series_bench(snap, steps: int, contexts: int = 2) -> dict:
    """The same steps as a series of snapshots on `contexts` device contexts
    at once (the reference's snapshot loop, src/main.cpp:56-79): one
    snapshot's host clump logic overlaps the next one's kernels.  Resident:
    each context restores its own device copy of the snapshot per step;
    with_upload: each step starts from the float32 snapshot in host memory.
    Every context runs one untimed step first."""
    import threading

    import plan_amd

    ctxs = [plan_amd.Context(0) for _ in range(contexts)]
    try:
        for c in ctxs:
            c.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
            c.keep_input()
        out = {"contexts": contexts, "snapshots": steps}
        for mode in ("resident", "with_upload"):
            def one(c):
                if mode == "resident":
                    c.restore_input()
                else:
                    c.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
                c.make_ghosts()
                c.build_tree(True)
                return len(c.find_planetesimals()["clumps"])

            for c in ctxs:
                one(c)
            lock = threading.Lock()
            todo = [steps]
            found = []
            errors = []
            stages = []

            def worker(c):
                try:
                    while True:
                        with lock:
                            if todo[0] == 0:
                                return
                            todo[0] -= 1
                        found.append(one(c))
                        st = c.stage_times()
                        cst = c.clump_stats()
                        for k in ("ms_clump_logic", "ms_sort", "ms_mergeA", "ms_rebind", "rebind_max_ms",
                                  "ms_merge_saddle", "rebind_exact_total_ms", "d2h_ms"):
                            st[k] = cst[k]
                        stages.append(st)
                except BaseException as e:  # reported after the join
                    errors.append(e)

            th = [threading.Thread(target=worker, args=(c,)) for c in ctxs]
            t0 = time.perf_counter()
            for t in th:
                t.start()
            for t in th:
                t.join()
            dt = time.perf_counter() - t0
            if errors:
                raise errors[0]
            out[mode] = {"value": snap.n * steps / dt, "unit": "particles/s", "ms_per_snapshot": dt / steps * 1e3,
                         "clumps_found": sorted(set(found)),
                         "stage_ms": {k: float(np.mean([x[k] for x in stages])) for k in stages[0]}}
            log(f"[bench] series {mode}: {out[mode]}")
        return out
    finally:
        for c in ctxs:
            c.close()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads: int, snap_full, c1_runs: int = 3) -> dict:
    """The oracle (CPU restatement of the reference, OpenMP) end to end:
    ghosts + BuildTree + FindPlanetesimals on the benchmark snapshot itself
    (one run, about 20 s on 16 cores), and on the reference's CPU-runnable
    262,144-particle configuration (median of c1_runs runs)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as O
    from plan_amd import synth

    O.set_threads(threads)
    c1 = synth.config_c1(seed=0)
    c1_s = []
    for _ in range(c1_runs):
        t0 = time.perf_counter()
        O.run_pipeline(c1)
        c1_s.append(time.perf_counter() - t0)
    c1_med = float(np.median(c1_s))
    t0 = time.perf_counter()
    t = O.run_pipeline(snap_full)
    dt = time.perf_counter() - t0
    ncl = len(t.catalog()["clumps"])
    del t
    return dict(value=snap_full.n / dt, unit="particles/s", cores=threads, kind="port",
                cpu=cpu_model(),
                sample=f"oracle end to end (ghosts+BuildTree+FindPlanetesimals) on the whole {snap_full.n}-particle "
                       f"benchmark snapshot: {dt:.2f} s, {ncl} clumps (one run)",
                c1={"n": c1.n, "median_s": c1_med, "runs_s": c1_s, "particles_per_s": c1.n / c1_med})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=None,
                    help="particles (default 16,777,216 for the plain workload, 134,217,728 for large)")
    ap.add_argument("--workload", choices=("plain", "large"), default="plain",
                    help="plain: 200 clumps in 10%% of the particles, K=32; large: 2,000 clumps in 15%%, K=64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-boundness", action="store_true")
    ap.add_argument("--boundness-clumps", type=int, default=5000)
    ap.add_argument("--no-series", action="store_true")
    ap.add_argument("--series-contexts", type=int, default=2)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import plan_amd
    from plan_amd import synth

    if args.n is None:
        args.n = 16777216 if args.workload == "plain" else 134217728
    if args.workload == "large":
        # one context holds about 108 GiB at this size: a second context for
        # the series would not fit, and the CPU restatement would run minutes
        args.no_series = True
        args.no_cpu_baseline = True
        args.no_boundness = True
    log(f"[bench] generating {args.workload} snapshot n={args.n}")
    t0 = time.perf_counter()
    snap = synth.config_c2(n=args.n) if args.workload == "plain" else synth.config_c5(n=args.n)
    log(f"[bench] generated in {time.perf_counter() - t0:.1f}s; uploading")
    ctx = plan_amd.Context(0)
    t0 = time.perf_counter()
    ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
    upload_s = time.perf_counter() - t0
    ctx.keep_input()

    def step():
        ctx.restore_input()
        T = ctx.make_ghosts()
        info = ctx.build_tree(True)
        cat = ctx.find_planetesimals()
        return T, info, cat

    for w in range(args.warmup):
        T, info, cat = step()
        log(f"[bench] warmup {w}: T={T} n={info.num_particles} nodes={info.num_nodes} clumps={len(cat['clumps'])} "
            f"stages={ctx.stage_times()}")
    kms = []
    rms = []
    stages = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        T, info, cat = step()
        kms.append(ctx.density_kernel_ms())
        rms.append(ctx.density_radius_ms())
        st = ctx.stage_times()
        stages.append(st)
        log(f"[bench] step {s}: stages={st} clump_stats={ctx.clump_stats()}")
    total = time.perf_counter() - t0
    ms_per_step = total / args.steps * 1e3
    value = args.n / (total / args.steps)
    # the same steps with the snapshot handed over from host memory each time
    # (float32 fields through PCIe, converted on the device): the
    # PCIe-inclusive rate, reported beside the resident-input value
    t0 = time.perf_counter()
    for s in range(args.steps):
        ctx.set_particles(snap.params, snap.pos, snap.vel, snap.density, snap.property_index, snap.id)
        ctx.make_ghosts()
        ctx.build_tree(True)
        ctx.find_planetesimals()
    up_total = time.perf_counter() - t0
    log(f"[bench] with upload: {up_total / args.steps * 1e3:.1f} ms per step")
    k_ms = float(np.mean(kms))
    nparticles_tree = info.num_particles
    # the timed launch is the density launch: one query per tree particle
    launch_units = nparticles_tree
    achieved = DENSITY_BYTES_PER_PARTICLE * launch_units / (k_ms * 1e-3) / 1e9
    traffic = measured_traffic(launch_units)
    mean_st = {k: float(np.mean([s[k] for s in stages])) for k in stages[0]}
    series = None
    if not args.no_series:
        series = series_bench(snap, args.steps, args.series_contexts)
    boundness = None
    if not args.no_boundness:
        boundness = boundness_bench(ctx, args, snap.params)
    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"[bench] cpu baseline on {threads} threads ({cpu_model()})")
        cpu = cpu_baseline(threads, snap)
        log(f"[bench] cpu baseline: {cpu}")
    out = {
        "metric": ("particles/s end-to-end clump finding @16.8M" if args.workload == "plain"
                   else "particles/s end-to-end clump finding @134M"),
        "value": value,
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"plain drop-in: {args.n} particles, 200 Plummer clumps (10%), box [-0.1,0.1]^3, K=32, "
                         if args.workload == "plain" else
                         f"large: {args.n} particles, 2,000 Plummer clumps (15%), box [-0.1,0.1]^3, K=64, ")
                        + "H=64, seed 0; ghosts+BuildTree+FindPlanetesimals per step",
            "N": args.n, "T_with_ghosts": T, "tree_particles": nparticles_tree, "nodes": info.num_nodes,
            "clumps_found": len(cat["clumps"]), "parallelism": "single GPU",
            "stage_ms": mean_st, "h2d_upload_s": upload_s,
        },
        "with_upload": {
            "value": args.n / (up_total / args.steps), "unit": "particles/s", "ms_per_step": up_total / args.steps * 1e3,
            "note": "each step starts from the float32 snapshot in host memory (plan_set_particles_f32 over PCIe)",
        },
        "series": series,
        "roofline": {
            "kernel": "k_knn_select<32,false> (density pass: walk, K-nearest selection, median density)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic,
            "units_per_launch": launch_units,
            "bytes_per_unit": DENSITY_BYTES_PER_PARTICLE,
            "kernel_ms": k_ms,
            "radius_kernel_ms": float(np.mean(rms)),
        },
        "cpu_baseline": cpu,
        "boundness": boundness,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
"""Benchmark of the pair-counting correlation estimator on one MI355X.

A step is one full estimator pass (tos_xi_run) over every candidate pixel
pair of the resident catalogue.  The default workload is the full-BOSS-scale
precision test: 160,000 synthetic Lyman-alpha quasars over 10,000 deg^2
(~2.3e7 pixels), HEALPix order 4 candidate search, Cartesian 50x50 grid over
0-200 Mpc/h.  `value` is the FP32C (FP32 separations, FP64 guard band)
throughput in pixel pairs per second; the FP64 kernel's throughput, their
ratio, the VALU roofline fraction, the p50/p99 latency of a short served
job stream (Poisson arrivals, mixed sizes, grids and precisions) and the CPU
oracle baseline (all host threads, bounded sample) are reported on the same
JSON line.

  --config c1   the same on 10,000 quasars over 1000 deg^2
  --config c3   batched mocks: 128 realisations in one hipGraph-captured
                batched launch against 128 sequential runs (own JSON line)
  --config c4   a served stream of --jobs requests at --qps (own JSON line)
  --config c0   3D brute-force xi of 10,000 points (bruteGPU / BinPair; own line)
  --config n4   brute-force sweep over --sweep point counts (own JSON line)
  --config n1   GPU candidate search against the host search (own JSON line)
  --config n2   per-sky-bin subsample pass and covariance (own JSON line)
  --config n3   per-pixel distance transform on the device (own JSON line)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c2|c1|c0|c3|c4|n1|n2|n3|n4]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "turbo-octo-spice_amd"))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

# published MI355X peaks (TFLOP/s)
PEAK_FP32 = 157.3
PEAK_FP64 = 78.6
# algorithmic flops per evaluated pixel pair (reference formula, once per pair)
F_ALG = {"cartesian": 20, "polar": 18, "observing": 11}

CONFIGS = {
    "c2": dict(n_qso=160000, area=10000.0, seed=3000,
               workload="full-BOSS-scale sightlines: 160000 QSOs / 10000 deg2, Cartesian 50x50 [0,200] Mpc/h"),
    "c1": dict(n_qso=10000, area=1000.0, seed=2000,
               workload="Lya sightlines: 10000 QSOs / 1000 deg2, Cartesian 50x50 [0,200] Mpc/h"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(cat, grid, target_s=15.0):
    """The CPU oracle (the reference algorithm restated in C, one task per
    occupied sky bin on a thread pool) timed on a bounded sample of the sky
    bins: every k-th task, k chosen so the run takes ~target_s.  Returns the
    baseline entry and the oracle's per-task histograms of that sample."""
    from oracle import oracle as ora
    nthreads = min(os.cpu_count() or 1, 16)
    stride = 64
    r = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=nthreads, task_stride=stride)
    t0 = time.perf_counter()
    r = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=nthreads, task_stride=stride, per_skybin=True)
    dt = time.perf_counter() - t0
    if dt < target_s / 2:
        stride = max(1, int(stride * dt / target_s))
        t0 = time.perf_counter()
        r = ora.xi_run(cat, grid, cat.scale, order=4, nthreads=nthreads, task_stride=stride, per_skybin=True)
        dt = time.perf_counter() - t0
    pp = r["stats"]["num_pixel_pairs"]
    return ({"value": pp / dt, "unit": "pixel-pairs/s", "cores": nthreads, "kind": "port",
             "sample": f"one in {stride} occupied HEALPix cell tasks ({r['n_tasks']} tasks, "
                       f"{pp:.3e} pixel pairs, {dt:.1f} s)"}, r)


def rel_err_max(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    m = (a != 0) | (b != 0)
    if not m.any():
        return 0.0
    return float(np.max(np.abs(a - b)[m] / np.maximum(np.abs(b[m]), 1e-300)))


def xi_of(didj, wgt):
    m = np.asarray(wgt) > 0
    return np.where(m, np.asarray(didj) / np.where(m, wgt, 1.0), 0.0)


def per_bin_error(h, ref):
    """Relative error of each bin's sums and xi of histogram h against ref
    (the FP64 pass): the largest over the bins, the median, and how many
    bins fall in each decade of error."""
    out = {"counts_equal": bool(np.array_equal(h.num_pairs, ref.num_pairs))}
    m = ref.wgt > 0
    for name, a, b in (("didj", h.didj, ref.didj), ("wgt", h.wgt, ref.wgt),
                       ("xi", xi_of(h.didj, h.wgt), xi_of(ref.didj, ref.wgt))):
        a = np.asarray(a)[m]
        b = np.asarray(b)[m]
        e = np.abs(a - b) / np.maximum(np.abs(b), 1e-300)
        out[f"max_rel_err_{name}"] = float(e.max()) if e.size else 0.0
        out[f"median_rel_err_{name}"] = float(np.median(e)) if e.size else 0.0
    e = np.abs(xi_of(h.didj, h.wgt)[m] - xi_of(ref.didj, ref.wgt)[m]) / np.maximum(np.abs(xi_of(ref.didj, ref.wgt)[m]), 1e-300)
    dec = {}
    for lo in range(-18, -5):
        dec[f"1e{lo}"] = int(((e >= 10.0 ** lo) & (e < 10.0 ** (lo + 1))).sum())
    dec["below"] = int((e < 1e-18).sum())
    dec["1e-6_or_more"] = int((e >= 1e-6).sum())
    out["bins_by_xi_rel_err_decade"] = dec
    out["bins"] = int(m.sum())
    return out


def sample_parity(pairs, grid, r):
    """GPU per-sky-bin histograms of the oracle's sampled tasks against the
    oracle's, both precisions: counts equal, largest relative sum error."""
    import tos_amd
    keys_ref = np.asarray(r["skybin_keys"])
    res = {"tasks": int(len(keys_ref))}
    for prec, name in ((tos_amd.FP64, "fp64"), (tos_amd.FP32C, "fp32c")):
        _, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=prec)
        idx = np.searchsorted(keys, keys_ref)
        ok = bool(np.array_equal(keys[np.minimum(idx, len(keys) - 1)], keys_ref))
        eq, ed, ew = ok, 0.0, 0.0
        if ok:
            for k, t in zip(idx, r["per_skybin"]):
                eq = eq and bool(np.array_equal(sn[k].astype(np.int64), t["num_pairs"]))
                ed = max(ed, rel_err_max(sd[k], t["didj"]))
                ew = max(ew, rel_err_max(sw[k], t["wgt"]))
        res[name] = {"counts_equal": eq, "max_rel_err_didj": ed, "max_rel_err_wgt": ew}
    return res


def copy_calibration(mib: int):
    import torch
    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 5}))


def traffic_from_profile(workload_key: str):
    """HBM bytes per launch of the FP32C pair kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (counters scaled by a
    plain copy of known size): its lines `traffic <config> fp32c <bytes>`.
    None when no summary covers the workload."""
    path = os.path.join(ROOT, "profiles", "xi_pairs_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) == 4 and tok[0] == "traffic" and tok[1] == workload_key and tok[2] == "fp32c":
                return float(tok[3])
    return None


def counters_from_profile(workload_key: str):
    """VALU issue rate and LDS busy fraction of the FP32C pair kernel on
    this workload, from the committed rocprofv3 SQ counter summary: its
    lines `counters <config> fp32c <valu per CU-cycle> <lds busy>`.  None
    when no summary covers the workload."""
    path = os.path.join(ROOT, "profiles", "xi_pairs_c2_sq_counters.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            tok = line.split()
            if len(tok) == 5 and tok[0] == "counters" and tok[1] == workload_key and tok[2] == "fp32c":
                return {"valu_instructions_per_cu_cycle": float(tok[3]), "lds_busy_of_cu_cycles": float(tok[4]),
                        "source": "profiles/xi_pairs_c2_sq_counters.txt (rocprofv3 SQ counters, peak 1 VALU "
                                  "instruction per CU-cycle)"}
    return None


def nthreads_host():
    return min(os.cpu_count() or 1, 16)


def bench_c3(args):
    """Batched mocks: 128 realisations of the Lyman-alpha generator (seeds
    4000+i, sightline counts uniform in [2000, 40000], 2000 deg^2), Cartesian
    50x50, FP32C, all in one hipGraph-captured batched launch with the
    autotuned launch shape; against 128 sequential runs."""
    import numpy as np
    import tos_amd
    from tos_amd import GridSpec, synth
    n_mocks = args.mocks
    rng = np.random.default_rng(4000)
    sizes = rng.integers(2000, 40001, n_mocks)
    grid = GridSpec.cartesian()
    dev = tos_amd.Device(0)
    t0 = time.perf_counter()
    mocks = []
    npp = 0
    for i, n in enumerate(sizes):
        cat = synth.lya_catalog(int(n), area_deg2=2000.0, seed=4000 + i)
        dcat = dev.upload(cat)
        pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=nthreads_host())
        npp += pairs.stats()["num_pixel_pairs"]
        mocks.append((dcat, pairs))
        del cat
    log(f"{n_mocks} mocks, {int(sizes.sum())} sightlines, {npp:.4e} pixel pairs "
        f"({time.perf_counter() - t0:.1f} s to generate and upload)")
    big = int(np.argmax(sizes))
    block, bpc, tms = dev.autotune(mocks[big][1], grid, tos_amd.FP32C)
    shape = dev.tuned_shape(grid, tos_amd.FP32C)
    log(f"autotuned launch shape on the largest mock ({int(sizes[big])} QSOs): {shape} ({tms:.2f} ms; "
        "0: the default)")
    jobs = [(p, grid, tos_amd.FP32C) for _, p in mocks]
    for _ in range(args.warmup):
        dev.batch_run(jobs, use_graph=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res, bms = dev.batch_run(jobs, use_graph=True)
    dt = time.perf_counter() - t0
    v_batch = npp * args.steps / dt
    log(f"batched: {args.steps} passes {dt:.3f} s, device {bms:.1f} ms per pass")
    t0 = time.perf_counter()
    seq = [p.run(grid, precision=tos_amd.FP32C) for _, p in mocks]
    dts = time.perf_counter() - t0
    same = all(np.array_equal(a.num_pairs, b.num_pairs) for a, b in zip(res, seq))
    v_seq = npp / dts
    log(f"sequential: {dts:.3f} s")
    out = {
        "metric": "pixel-pairs/s, FP32-comp batched mocks on 1 GPU",
        "value": v_batch, "unit": "pixel-pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded generator of the configuration; no survey data)",
        "config": {"workload": f"batched mocks: {n_mocks} realisations, 2000-40000 QSOs each over 2000 deg2, "
                               "Cartesian 50x50 [0,200] Mpc/h, FP32C, one hipGraph-captured batched launch",
                   "n_mocks": n_mocks, "n_qso_total": int(sizes.sum()), "pixel_pairs_per_step": npp,
                   "autotuned": shape},
        "sequential_value": v_seq,
        "batched_over_sequential": v_batch / v_seq,
        "counts_batched_equal_sequential": bool(same),
        "device_ms_per_batched_pass": bms,
    }
    for d, p in mocks:
        p.free()
        d.free()
    dev.close()
    print(json.dumps(out))


def served_stream(dev, n_jobs, qps, seed=5000, max_qso=100000, area=1000.0, search="device"):
    """A served stream: Poisson arrivals at `qps`, sightline counts
    log-uniform in [1000, max_qso] over `area` deg^2, half Cartesian 50x50 /
    half Polar 50x20, half FP64 / half FP32C.  Catalogues are generated
    before the clock starts (the client's data); returns a summary dict."""
    import threading

    import numpy as np
    import tos_amd
    from tos_amd import GridSpec, synth
    from tos_amd.server import Server
    rng = np.random.default_rng(seed)
    grids = [GridSpec.cartesian(), GridSpec.polar()]
    precs = [tos_amd.FP64, tos_amd.FP32C]
    reqs = []
    t0 = time.perf_counter()
    for k in range(n_jobs):
        n = int(np.exp(rng.uniform(np.log(1000), np.log(max_qso))))
        cat = synth.lya_catalog(n, area_deg2=area, seed=seed + 1 + k)
        reqs.append((cat, grids[int(rng.integers(2))], precs[int(rng.integers(2))]))
    gaps = rng.exponential(1.0 / qps, n_jobs)
    log(f"served stream: {n_jobs} jobs generated in {time.perf_counter() - t0:.1f} s, "
        f"{sum(c.n_los for c, _, _ in reqs)} sightlines, QPS {qps}")
    ids = [None] * n_jobs
    with Server(dev, hbm_budget_bytes=96 << 30, max_inflight=64, nthreads=nthreads_host(),
                search_on_device=search == "device") as srv:
        def client():
            for k, (cat, g, p) in enumerate(reqs):
                time.sleep(gaps[k])
                ids[k] = srv.submit(cat, g, p)
        th = threading.Thread(target=client)
        th.start()
        th.join()
        timings, npp, ok = [], 0, 0
        for k in range(n_jobs):
            srv.wait(ids[k], 600000)
            st, h, t = srv.result(ids[k])
            # device time of the launches that carried the job's slices
            t["launch_device_ms"] = float(h.kernel_ms) if h is not None else 0.0
            timings.append(t)
            if st == 0:
                ok += 1
                npp += t["num_pixel_pairs"]
    lat = np.array([t["done_ms"] - t["submit_ms"] for t in timings])
    span = (max(t["done_ms"] for t in timings) - min(t["submit_ms"] for t in timings)) * 1e-3
    # each job alone on the device (the direct path): its slowdown in the
    # stream is its served latency over this isolated pass
    iso = []
    for cat, g, p in reqs:
        dcat = dev.upload(cat)
        pairs = dcat.pairs(g.max_angular_scale(cat.scale), order=4, nthreads=nthreads_host(), on_device=True)
        t0 = time.perf_counter()
        pairs.run(g, precision=p)
        iso.append((time.perf_counter() - t0) * 1e3)
        pairs.free()
        dcat.free()
    slow = lat / np.maximum(np.array(iso), 1e-3)

    def phases(t):
        # search + admission wait, wait for a launch slot, on the device
        return {"search_and_admit_ms": t["admit_ms"] - t["submit_ms"], "slot_wait_ms": t["start_ms"] - t["admit_ms"],
                "run_ms": t["done_ms"] - t["start_ms"], "launch_device_ms": t["launch_device_ms"],
                "pixel_pairs": int(t["num_pixel_pairs"])}
    slowest = phases(timings[int(np.argmax(lat))])
    # the job at the 99th percentile of slowdown, with its isolated time
    k99 = int(np.argsort(slow)[min(n_jobs - 1, int(round(0.99 * (n_jobs - 1))))])
    p99_job = dict(phases(timings[k99]), isolated_ms=float(iso[k99]), slowdown=float(slow[k99]),
                   sightlines=int(reqs[k99][0].n_los))
    med = {k: float(np.median([phases(t)[k] for t in timings])) for k in ("search_and_admit_ms", "slot_wait_ms",
                                                                           "run_ms")}
    return {"jobs": n_jobs, "completed": ok, "qps": qps,
            "slowest_job": slowest, "p99_slowdown_job": p99_job, "median_phases_ms": med,
            "p50_latency_ms": float(np.percentile(lat, 50)), "p99_latency_ms": float(np.percentile(lat, 99)),
            "slowdown_p50": float(np.percentile(slow, 50)), "slowdown_p99": float(np.percentile(slow, 99)),
            "slowdown": "served latency over the job's own pass alone on the device (tos_xi_run, "
                        "resident inputs)",
            "median_isolated_ms": float(np.median(iso)),
            "mean_latency_ms": float(lat.mean()), "pixel_pairs": int(npp),
            "throughput_pixel_pairs_per_s": npp / span if span > 0 else None,
            "sightlines": f"log-uniform [1000, {max_qso}] over {area:g} deg2", "candidate_search": search,
            "mix": "Cartesian 50x50 / Polar 50x20 x FP64 / FP32C, 50/50 each"}


def bench_n3(args):
    """The per-pixel distance transform (h5healxi.cc:196-216) of the
    full-survey catalogue's pixels on the device (tos_pixel_distances,
    host arrays in and out) against the host transform of the distinct
    wavelengths (numpy unique + tos_cosmo_distances + gather)."""
    import tos_amd
    from tos_amd import synth
    cfg = CONFIGS["c2"]
    cat = synth.lya_catalog(cfg["n_qso"], area_deg2=cfg["area"], seed=cfg["seed"])
    dev = tos_amd.Device(0)
    loglam = np.ascontiguousarray(cat.loglam, np.float32)
    times = []
    for _ in range(args.warmup + args.steps):
        t = time.perf_counter()
        d = dev.pixel_distances(loglam, 0.7, 0.3)
        times.append(time.perf_counter() - t)
    dev.close()
    t = time.perf_counter()
    uniq, inv = np.unique(loglam, return_inverse=True)
    dc, _ = synth.cosmo_distances(10.0 ** (uniq.astype(np.float64) - synth.LOG_LYA) - 1.0, 0.7, 0.3)
    host = dc.astype(np.float32)[inv]
    t_host = time.perf_counter() - t
    ulp = np.abs(d.view(np.int32).astype(np.int64) - host.view(np.int32).astype(np.int64))
    ms = float(np.median(times[args.warmup:])) * 1e3
    out = {"metric": "pixels/s, per-pixel FP64 distance transform on 1 GPU", "value": len(loglam) / ms * 1e3,
           "unit": "pixels/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded generator of the configuration; no survey data)",
           "config": {"workload": f"distance transform of the full-survey catalogue: {len(loglam)} pixels, "
                                  "512-point Gauss-Legendre integral per pixel, host arrays in and out"},
           "host_distinct_transform_ms": t_host * 1e3, "distinct_wavelengths": int(len(uniq)),
           "max_ulp_vs_host": int(ulp.max()), "fraction_equal": float((ulp == 0).mean())}
    print(json.dumps(out))


def bench_n1(args):
    """Candidate search (XiEstimator.cc:121-159) on the full-survey catalogue:
    the GPU search (tos_pairs_build, on_device = 1) against the host search on
    all host threads; both give the same pair list."""
    import tos_amd
    from tos_amd import GridSpec, synth
    cfg = CONFIGS["c2"]
    cat = synth.lya_catalog(cfg["n_qso"], area_deg2=cfg["area"], seed=cfg["seed"])
    grid = GridSpec.cartesian()
    dev = tos_amd.Device(0)
    dcat = dev.upload(cat)
    ma = grid.max_angular_scale(cat.scale)
    times = {False: [], True: []}
    for _ in range(args.warmup + args.steps):
        for on_dev in (False, True):
            t = time.perf_counter()
            p = dcat.pairs(ma, order=4, nthreads=nthreads_host(), on_device=on_dev)
            times[on_dev].append(time.perf_counter() - t)
            st = p.stats()
            p.free()
    dev_s = float(np.median(times[True][args.warmup:]))
    host_s = float(np.median(times[False][args.warmup:]))
    n = st["num_sightline_pairs"]
    log(f"candidate search: device {dev_s * 1e3:.1f} ms, host ({nthreads_host()} threads) {host_s * 1e3:.1f} ms, "
        f"{n} sightline pairs considered, {st['num_sightline_pairs_used']} kept")
    # the same with plate sky bins (a 2.5 deg tiling of the patch)
    plates = synth.plate_tiling(cat, spacing_deg=2.5)
    ptimes = {False: [], True: []}
    for _ in range(args.warmup + 1):
        for on_dev in (False, True):
            t = time.perf_counter()
            p = dcat.pairs(ma, plates=plates, nthreads=nthreads_host(), on_device=on_dev)
            ptimes[on_dev].append(time.perf_counter() - t)
            pst = p.stats()
            p.free()
    plate_line = {"n_plates": int(len(plates[1])), "device_ms": ptimes[True][-1] * 1e3,
                  "host_ms": ptimes[False][-1] * 1e3, "sightline_pairs_kept": pst["num_sightline_pairs_used"]}
    log(f"plate search ({len(plates[1])} plates): device {plate_line['device_ms']:.1f} ms, "
        f"host {plate_line['host_ms']:.1f} ms")
    dcat.free()
    dev.close()
    print(json.dumps({
        "metric": "candidate sightline pairs/s of the GPU candidate search on 1 GPU", "value": n / dev_s,
        "unit": "sightline-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dev_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded generator of the configuration; no survey data)",
        "config": {"workload": CONFIGS["c2"]["workload"] + ", HEALPix order 4 candidate search",
                   "sightline_pairs_considered": n, "sightline_pairs_kept": st["num_sightline_pairs_used"]},
        "host_search_ms": host_s * 1e3, "host_threads": nthreads_host(), "device_over_host": host_s / dev_s,
        "plates": plate_line}))


def bench_n2(args):
    """Per-sky-bin subsamples and their covariance on the full-survey
    catalogue: one per_skybin pass (every occupied sky bin a job of one
    batched launch) against the plain pass, then the subsample covariance."""
    import tos_amd
    from tos_amd import GridSpec, synth
    cfg = CONFIGS["c2"]
    cat = synth.lya_catalog(cfg["n_qso"], area_deg2=cfg["area"], seed=cfg["seed"])
    grid = GridSpec.cartesian()
    dev = tos_amd.Device(0)
    dcat = dev.upload(cat)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=nthreads_host())
    npp = pairs.stats()["num_pixel_pairs"]
    for _ in range(args.warmup):
        pairs.run_per_skybin(grid, precision=tos_amd.FP32C)
    t = time.perf_counter()
    for _ in range(args.steps):
        tot, keys, sn, sd, sw = pairs.run_per_skybin(grid, precision=tos_amd.FP32C)
    sub_s = (time.perf_counter() - t) / args.steps
    plain = pairs.run(grid, precision=tos_amd.FP32C)
    same = bool(np.array_equal(plain.num_pairs, tot.num_pairs))
    t = time.perf_counter()
    plain = pairs.run(grid, precision=tos_amd.FP32C)
    plain_s = time.perf_counter() - t
    dev.subsample_covariance(sd, sw)
    t = time.perf_counter()
    cov, mean, W = dev.subsample_covariance(sd, sw)
    cov_s = time.perf_counter() - t
    k, nb = sd.shape
    log(f"per-sky-bin pass {sub_s * 1e3:.1f} ms ({k} subsamples) against the plain pass {plain_s * 1e3:.1f} ms; "
        f"covariance of {k} x {nb} in {cov_s * 1e3:.1f} ms")
    pairs.free()
    dcat.free()
    dev.close()
    print(json.dumps({
        "metric": "pixel-pairs/s of the per-sky-bin subsample pass (FP32C) on 1 GPU", "value": npp / sub_s,
        "unit": "pixel-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": sub_s * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded generator of the configuration; no survey data)",
        "config": {"workload": CONFIGS["c2"]["workload"] + ", one histogram per occupied sky bin",
                   "n_subsamples": k, "nbins": nb},
        "plain_pass_ms": plain_s * 1e3, "totals_equal_plain_pass": same,
        "covariance_ms": cov_s * 1e3, "covariance_flop": 2.0 * k * nb * nb}))


def bench_c4(args):
    import tos_amd
    dev = tos_amd.Device(0)
    s = served_stream(dev, args.jobs, args.qps, search=args.server_search)
    dev.close()
    out = {"metric": "p50/p99 xi-job latency of a served stream on 1 GPU",
           "value": s["p99_latency_ms"], "unit": "ms (p99 latency)", "n_gpus": 1, "steps": args.jobs,
           "warmup": 0, "ms_per_step": s["mean_latency_ms"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32/f64 mix",
           "data": "synthetic (seeded generator of the configuration; no survey data)",
           "config": {"workload": f"served xi-job stream: {args.jobs} Poisson arrivals at {args.qps} QPS, "
                                  "1000-100000 QSOs log-uniform over 1000 deg2, Cartesian/Polar x FP64/FP32C"},
           "serving": s}
    print(json.dumps(out))


def c0_points(n, seed=1000):
    """Brute-force points: n positions uniform in a 100 Mpc/h cube, delta ~ N(0, 1),
    weight 1 (seeded)."""
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(0.0, 100.0, n) for _ in range(3))
    return x, y, z, rng.normal(0.0, 1.0, n), np.ones(n)


def bench_c0(args):
    """3D isotropic brute force (bruteGPU / BinPair): all N(N-1)/2 pairs of
    10,000 points, 50 bins over [0, 100) Mpc/h.  `value` is the FP64
    (BinPair) kernel's pairs per second with the points resident; the FP32
    (brute.cu semantics) kernel is reported beside it."""
    import tos_amd
    n, nb = 10000, 50
    pts = c0_points(n)
    pairs = n * (n - 1) // 2
    dev = tos_amd.Device(0)
    res = {}
    for mode, name in ((tos_amd.BRUTE_FP64, "fp64"), (tos_amd.BRUTE_FP32, "fp32")):
        for _ in range(args.warmup):
            dev.brute3d(*pts, 0.0, 100.0, nb, mode=mode)
        ks, walls = [], []
        for _ in range(args.steps):
            t = time.perf_counter()
            _, _, cn, ms = dev.brute3d(*pts, 0.0, 100.0, nb, mode=mode)
            walls.append(time.perf_counter() - t)
            ks.append(ms)
        res[name] = dict(kernel_ms=float(np.mean(ks)), call_ms=float(np.mean(walls)) * 1e3,
                         in_range=int(cn.sum()))
        log(f"brute {name}: kernel {res[name]['kernel_ms']:.3f} ms, call (upload + tiles + kernel + copy) "
            f"{res[name]['call_ms']:.3f} ms")
    dev.close()
    f64 = res["fp64"]
    achieved = pairs * 16 / (f64["kernel_ms"] * 1e-3) / 1e12
    out = {"metric": "pairs/s of the 3D brute-force xi on 1 GPU", "value": pairs / (f64["kernel_ms"] * 1e-3),
           "unit": "pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": f64["kernel_ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "f64", "data": "synthetic (seeded: uniform positions, N(0,1) deltas, unit weights)",
           "config": {"workload": "3D brute-force xi: 10000 points in a 100 Mpc/h cube, 50 bins [0,100) Mpc/h",
                      "n_points": n, "pairs_per_step": pairs},
           "fp32": {"value": pairs / (res["fp32"]["kernel_ms"] * 1e-3), "kernel_ms": res["fp32"]["kernel_ms"],
                    "roofline_frac_fp32": pairs * 16 / (res["fp32"]["kernel_ms"] * 1e-3) / 1e12 / PEAK_FP32},
           "host_inclusive_pairs_per_s": pairs / (f64["call_ms"] * 1e-3),
           "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64, "unit": "TFLOP/s",
                        "frac": achieved / PEAK_FP64, "traffic": None,
                        "compute_unit": "VALU FP64 (no MFMA); F_alg = 16 flop per pair",
                        "kernel_ms": f64["kernel_ms"]}}
    if not args.no_cpu:
        from oracle import oracle as ora
        ora.lib()
        t = time.perf_counter()
        ora.brute3d_f64(*pts, 0.0, 100.0, nb)
        dt = time.perf_counter() - t
        out["cpu_baseline"] = {"value": pairs / dt, "unit": "pairs/s", "cores": 1, "kind": "port",
                               "sample": f"the whole workload ({pairs} pairs, {dt:.2f} s, one thread)"}
    print(json.dumps(out))


# the reference's only published timings (README plot, digitised; the
# unlabeled fastest GPU curve): points -> seconds
PUBLISHED_BRUTE = {160000: 17.0, 2850000: 5300.0}


def bench_n4(args):
    """Large-N brute-force sweep over the published plot's range."""
    import tos_amd
    dev = tos_amd.Device(0)
    dev.brute3d(*c0_points(10000), 0.0, 100.0, 50, mode=tos_amd.BRUTE_FP32)   # warm-up
    rows = []
    for n in args.sweep:
        pts = c0_points(n, seed=1000 + n)
        pairs = n * (n - 1) / 2
        for mode, name in ((tos_amd.BRUTE_FP32, "fp32"), (tos_amd.BRUTE_FP64, "fp64")):
            t = time.perf_counter()
            _, _, cn, ms = dev.brute3d(*pts, 0.0, 100.0, 50, mode=mode)
            wall = time.perf_counter() - t
            row = {"n": n, "mode": name, "kernel_s": ms * 1e-3, "call_s": wall, "pairs_per_s": pairs / (ms * 1e-3)}
            if n in PUBLISHED_BRUTE:
                row["published_gpu_s"] = PUBLISHED_BRUTE[n]
            rows.append(row)
            log(f"n {n} {name}: kernel {ms * 1e-3:.4f} s, {pairs / (ms * 1e-3):.3e} pairs/s")
    dev.close()
    big = [r for r in rows if r["mode"] == "fp32"][-1]
    out = {"metric": "pairs/s of the 3D brute-force xi, large-N sweep on 1 GPU", "value": big["pairs_per_s"],
           "unit": "pairs/s", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": big["kernel_s"] * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded: uniform positions in a 100 Mpc/h cube, N(0,1) deltas, unit weights)",
           "config": {"workload": f"3D brute-force sweep, N = {args.sweep}, 50 bins [0,100) Mpc/h; value at "
                                  f"N = {big['n']} (FP32, the reference GPU kernel's semantics)"},
           "sweep": rows}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS) + ["c0", "c3", "c4", "n1", "n2", "n3", "n4"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--mocks", type=int, default=128, help="c3: number of mock realisations")
    ap.add_argument("--jobs", type=int, default=64, help="c4: served jobs")
    ap.add_argument("--qps", type=float, default=1.0, help="c4: Poisson arrival rate")
    ap.add_argument("--sweep", type=int, nargs="+", default=[4000, 40000, 160000, 1000000, 2850000],
                    help="n4: point counts")
    ap.add_argument("--server-search", default="device", choices=["device", "host"],
                    help="served streams: candidate search on the GPU or on the host threads")
    ap.add_argument("--serve-jobs", type=int, default=64,
                    help="served-stream jobs measured beside the default line (0: none)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB five times (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); the estimator is a "
              "single-device build (several GPUs would each take a share of the sky-bin tasks "
              "and sum their histograms)", file=sys.stderr)
        sys.exit(2)

    if args.config == "c0":
        bench_c0(args)
        return
    if args.config == "n4":
        bench_n4(args)
        return
    if args.config == "n3":
        return bench_n3(args)
    if args.config == "n1":
        bench_n1(args)
        return
    if args.config == "n2":
        bench_n2(args)
        return
    if args.config == "c3":
        bench_c3(args)
        return
    if args.config == "c4":
        bench_c4(args)
        return

    import tos_amd
    from tos_amd import GridSpec, synth

    cfg = CONFIGS[args.config]
    t0 = time.perf_counter()
    cat = synth.lya_catalog(cfg["n_qso"], area_deg2=cfg["area"], seed=cfg["seed"])
    log(f"generated {cat.n_los} sightlines, {cat.n_pix} pixels in {time.perf_counter() - t0:.1f} s")
    grid = GridSpec.cartesian()
    polar = GridSpec.polar()
    dev = tos_amd.Device(0)
    t0 = time.perf_counter()
    dcat = dev.upload(cat)
    # Cartesian 50x50 and Polar 50x20 have the same angular scale (both
    # reach 200 Mpc/h transverse), so one candidate pair list serves both
    assert grid.max_angular_scale(cat.scale) == polar.max_angular_scale(cat.scale)
    pairs = dcat.pairs(grid.max_angular_scale(cat.scale), order=4, nthreads=nthreads_host(), on_device=True)
    st = pairs.stats()
    log(f"candidate pairs: {st['num_sightline_pairs_used']} sightline pairs, "
        f"{st['num_pixel_pairs']:.4e} pixel pairs ({time.perf_counter() - t0:.1f} s)")
    npp = st["num_pixel_pairs"]

    def timed(g, prec, steps, warmup):
        for _ in range(warmup):
            h = pairs.run(g, precision=prec)
        kms = []
        t0 = time.perf_counter()
        for _ in range(steps):
            h = pairs.run(g, precision=prec)
            kms.append(h.kernel_ms)
        dt = time.perf_counter() - t0
        log(f"grid {g.type} precision {prec}: {steps} steps {dt:.3f} s, kernel avg {np.mean(kms):.3f} ms, "
            f"rechecks {h.num_guard_rechecks}, evaluated {h.num_pairs_evaluated:.4e}")
        return dt, float(np.mean(kms)), h

    n64 = args.fp64_steps if args.fp64_steps is not None else max(1, args.steps // 2)

    def both(g, fname, steps):
        dt32, k32, h32 = timed(g, tos_amd.FP32C, steps, args.warmup)
        dt64, k64, h64 = timed(g, tos_amd.FP64, n64, min(args.warmup, 1))
        f = F_ALG[fname]
        return {"dt32": dt32, "k32": k32, "h32": h32, "dt64": dt64, "k64": k64, "h64": h64,
                "v32": npp * steps / dt32, "v64": npp * n64 / dt64,
                "achieved": h32.num_pairs_evaluated * f / (k32 * 1e-3) / 1e12,
                "achieved64": h64.num_pairs_evaluated * f / (k64 * 1e-3) / 1e12}

    C = both(grid, "cartesian", args.steps)
    P = both(polar, "polar", max(1, args.steps // 2))
    out = {
        "metric": "pixel-pairs/s, FP32-comp vs plain FP64 on 1 GPU; p50/p99 xi-job latency",
        "value": C["v32"],
        "unit": "pixel-pairs/s",
        "value_counts": "candidate pixel pairs per second: sum of n_i n_j over the sightline pairs that pass "
                        "the angular cut (XiEstimator.cc:149 num_pixel_pairs); the kernel evaluates "
                        "pixel_pairs_evaluated_per_step of them, the rest lie outside the r_par/r_perp "
                        "windows and are rejected by the reference arithmetic too",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": C["dt32"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator of the configuration; no survey data)",
        "config": {"workload": cfg["workload"], "n_qso": cfg["n_qso"], "area_deg2": cfg["area"],
                   "n_pixels": cat.n_pix, "healpix_order": 4,
                   "grid": "cartesian [0:200]*50 x [0:200]*50 x [3.556:3.735]*1",
                   "pixel_pairs_per_step": npp, "precision": "fp32c"},
        "pixel_pairs_per_step": npp,
        "pixel_pairs_evaluated_per_step": C["h32"].num_pairs_evaluated,
        "fp64": {"value": C["v64"], "ms_per_step": C["dt64"] / n64 * 1e3, "steps": n64, "kernel_ms": C["k64"],
                 "roofline_frac_fp64": C["achieved64"] / PEAK_FP64},
        "fp32c_over_fp64": C["v32"] / C["v64"],
        "kernel_fp32c_over_fp64": C["k64"] / C["k32"],
        "counts_fp32c_equal_fp64": bool(np.array_equal(C["h32"].num_pairs, C["h64"].num_pairs)),
        "fp32c_per_bin_error_vs_fp64": per_bin_error(C["h32"], C["h64"]),
        "guard_rechecks_per_pair": C["h32"].num_guard_rechecks / max(1, C["h32"].num_pixel_pairs_used),
        "roofline": {"bound": "valu", "achieved": C["achieved"], "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": C["achieved"] / PEAK_FP32, "traffic": traffic_from_profile(args.config),
                     "compute_unit": "VALU (no MFMA: pair counting is not a dense contraction); "
                                     "peak = MI355X FP32 157.3 TFLOP/s; F_alg = 20 flop per evaluated "
                                     "pixel pair",
                     "frac_of_unpacked_fp32": C["achieved"] / (PEAK_FP32 / 2),
                     "achieved_over_all_candidate_pairs": npp * F_ALG["cartesian"] / (C["k32"] * 1e-3) / 1e12,
                     "kernel_ms": C["k32"], "issue": counters_from_profile(args.config)},
        "polar": {"grid": "polar [0:200]*50 x [0:1]*20 x [3.556:3.735]*1",
                  "value_fp32c": P["v32"], "value_fp64": P["v64"], "fp32c_over_fp64": P["v32"] / P["v64"],
                  "kernel_ms_fp32c": P["k32"], "kernel_ms_fp64": P["k64"],
                  "pixel_pairs_evaluated_per_step": P["h32"].num_pairs_evaluated,
                  "counts_fp32c_equal_fp64": bool(np.array_equal(P["h32"].num_pairs, P["h64"].num_pairs)),
                  "fp32c_per_bin_error_vs_fp64": per_bin_error(P["h32"], P["h64"]),
                  "roofline": {"bound": "valu", "achieved": P["achieved"], "peak": PEAK_FP32, "unit": "TFLOP/s",
                               "frac": P["achieved"] / PEAK_FP32, "f_alg": F_ALG["polar"],
                               "frac_fp64": P["achieved64"] / PEAK_FP64}},
    }
    if not args.no_cpu:
        out["cpu_baseline"], sample = cpu_baseline(cat, grid)
        out["oracle_sample_parity"] = sample_parity(pairs, grid, sample)
    pairs.free()
    dcat.free()
    if args.serve_jobs > 0:
        out["serving"] = served_stream(dev, args.serve_jobs, 2.0, search=args.server_search)
    dev.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark: gene x target Mann-Whitney U tests/sec on one MI355X.

Default workload: 150,000 cells x 8,000 genes of integer UMI counts (float32),
reference 15,000 cells, 300 targets, zero_handling='min', tie + continuity
correction, asymptotic p.  One step = one pass of the batched form (the launch
sequence over the resident matrix, columns largest first, replayed from a
hipGraph), producing all 2.4M (gene, target) U1 and p-values.

Reported beside it (extra keys): the plain per-call form (one C-ABI call per
job: upload, the same kernels launched eagerly, download) and its device-only
time; the roofline of the dominant kernel, timed with HIP events around it; a
CPU baseline (the oracle, a C restatement of the reference algorithm, timed on
a bounded column sample with the host's cores).

python bench.py [--steps K] [--warmup W] [--gpus 1] [--workload histogram]
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


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


_WORKLOAD_ARRAYS = ("data", "indices", "indptr", "group_id")


def generate_workload(name):
    from hpdex_amd import synthetic
    if name == "histogram":
        return synthetic.histogram_case()
    if name == "small":
        return synthetic.small_case()
    if name == "float":
        return synthetic.float_case()
    if name == "float-mix":
        return synthetic.float_case(zero_handling="mix")
    if name == "exact":
        return synthetic.exact_case()
    raise SystemExit(f"unknown workload {name}")


def load_workload(name, cache=None):
    """The seeded workload; with `cache` (a directory) its arrays are read
    from .npy files there when a previous run saved them (the generators are
    deterministic, so this only saves the generation time of repeated runs,
    e.g. one per profiler pass), else generated and saved there."""
    if not cache:
        return generate_workload(name)
    d = os.path.join(cache, name)
    meta = os.path.join(d, "meta.json")
    if os.path.exists(meta):
        with open(meta) as f:
            w = json.load(f)
        for k in _WORKLOAD_ARRAYS:
            w[k] = np.load(os.path.join(d, k + ".npy"))
        return w
    w = generate_workload(name)
    os.makedirs(d, exist_ok=True)
    for k in _WORKLOAD_ARRAYS:
        np.save(os.path.join(d, k + ".npy"), w[k])
    with open(meta + ".tmp", "w") as f:
        json.dump({k: v for k, v in w.items() if k not in _WORKLOAD_ARRAYS}, f)
    os.replace(meta + ".tmp", meta)
    return w


def traffic_from_profiles(kernel, workload):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE, each scaled by a streaming copy of known size),
    as recorded in profiles/<kernel>_<workload>_traffic.json; None if absent."""
    path = os.path.join(ROOT, "profiles", f"{kernel}_{workload}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    return d.get("bytes_per_launch")


def host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(w, target_s=12.0):
    """The oracle (C restatement of the reference algorithm, OpenMP) on the
    host cores: the whole workload when one pass is short (median of >= 3
    timed passes after a warm-up, about target_s in total), else an evenly
    spaced column sample sized to about target_s."""
    import oracle
    from hpdex_amd import synthetic
    threads = min(16, os.cpu_count() or 1)
    o = w["opts"]
    C = w["n_cols"]

    def run(sub):
        t0 = time.perf_counter()
        oracle.mwu(sub["data"], sub["indices"], sub["indptr"], sub["group_id"], sub["n_targets"],
                   n_rows=sub["n_rows"], tie_correction=o["tie_correction"],
                   use_continuity=o["use_continuity"], zero_handling=o["zero_handling"],
                   method=o["method"], threads=threads)
        return time.perf_counter() - t0

    # calibrate on a stride sample
    stride = max(1, C // 100)
    cols = np.arange(0, C, stride)
    dt = run(synthetic.subset_columns(w, cols))
    est_full = dt * C / len(cols)
    if est_full <= target_s / 3:
        sub = w
        run(sub)                                   # warm-up
        reps = int(min(50, max(3, target_s / max(est_full, 1e-3))))
        times = sorted(run(sub) for _ in range(reps))
        dt = times[len(times) // 2]
        ncols = C
        how = f"whole workload, median of {reps} passes after a warm-up ({dt:.3f} s per pass)"
    else:
        ncols = int(min(C, max(len(cols), len(cols) * target_s / max(dt, 1e-3))))
        cols = np.unique(np.linspace(0, C - 1, ncols).astype(np.int64))
        dt = run(synthetic.subset_columns(w, cols))
        ncols = len(cols)
        how = f"{ncols} of {C} genes (evenly spaced), one pass of {dt:.1f} s"
    tests = ncols * w["n_targets"]
    return {"value": tests / dt, "unit": "tests/s", "cores": threads, "kind": "port",
            "sample": f"{how}; oracle C restatement of the reference algorithm (quicksort "
                      f"merges, OpenMP over genes); host {host_cpu()}, {os.cpu_count()} CPUs "
                      f"visible, {threads} threads used"}


def ingest_bench(w, reps=3):
    """The matrix as AnnData holds it -- CSR, int32 indices -- made resident
    two ways: the host route the reference's driver and wrapper take
    (scipy tocsc, int64 copies of indices and indptr, then the CSC upload)
    and the device route (hpdex_matrix_create_from: CSR upload, transpose on
    the device).  Median wall time of `reps` runs each, after one warm-up."""
    import scipy.sparse as sp
    from hpdex_amd import _capi
    R, C = w["n_rows"], w["n_cols"]
    csr = sp.csc_matrix((w["data"], w["indices"], w["indptr"]), shape=(R, C)).tocsr()
    csr.indices = csr.indices.astype(np.int32)
    csr.indptr = csr.indptr.astype(np.int32)

    def host_route():
        t0 = time.perf_counter()
        h = csr.tocsc()
        ind, ptr = h.indices.astype(np.int64), h.indptr.astype(np.int64)
        t1 = time.perf_counter()
        dm = _capi.DeviceMatrix(h.data, ind, ptr, R)
        t2 = time.perf_counter()
        dm.close()
        return t1 - t0, t2 - t1

    def device_route():
        t0 = time.perf_counter()
        dm = _capi.DeviceMatrix.from_scipy(csr)
        t1 = time.perf_counter()
        dm.close()
        return t1 - t0

    device_route()
    dev = sorted(device_route() for _ in range(reps))[reps // 2]
    host = sorted(host_route() for _ in range(reps))
    conv = sorted(h[0] for h in host)[reps // 2]
    up = sorted(h[1] for h in host)[reps // 2]
    return {"input": f"CSR {R} x {C}, nnz {csr.nnz}, float32 values, int32 indices",
            "host_tocsc_and_int64_ms": conv * 1e3, "host_csc_upload_ms": up * 1e3,
            "host_route_ms": (conv + up) * 1e3, "device_route_ms": dev * 1e3,
            "host_over_device": (conv + up) / dev}


def stream_bench(w, limit_gb, workers=3):
    """The workload's DE (hpdex_amd.stream) streamed from memory-mapped .npy
    files on local disk under a memory limit, against one hpdex_de_csc call
    on the arrays in host memory; identical outputs are checked."""
    import tempfile
    from hpdex_amd import _capi, stream
    R, T = w["n_rows"], w["n_targets"]
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        arrs = []
        for name in ("data", "indices", "indptr"):
            f = os.path.join(d, name + ".npy")
            np.save(f, w[name])
            arrs.append(np.load(f, mmap_mode="r"))
        runs = stream.plan_gene_runs(arrs[2], T, int(limit_gb * 2**30) // workers, R,
                                     arrs[0].dtype)
        t0 = time.perf_counter()
        got = stream.stream_de(*arrs, R, w["group_id"], T, memory_limit_gb=limit_gb,
                               num_workers=workers)
        ts = time.perf_counter() - t0
    t0 = time.perf_counter()
    want = _capi.de_csc(w["data"], w["indices"], w["indptr"], w["group_id"], T, n_rows=R)
    to = time.perf_counter() - t0
    same = all(np.array_equal(g, x, equal_nan=True) for g, x in zip(got, want))
    return {"memory_limit_gb": limit_gb, "workers": workers, "runs": len(runs),
            "stream_ms": ts * 1e3, "one_shot_ms": to * 1e3, "identical": same,
            "source": "memory-mapped .npy files (data, int64 indices, indptr) on local disk"}


def serving_bench(n_jobs=64, rate=8.0, slots=4, budget_gb=48, seed=1005):
    """Serving workload (cells U[5k,200k], genes U[2k,20k], density U[3%,12%],
    NB(LogNormal(0,1), r=0.7)+1 counts, reference 10% of cells, 5-500 targets;
    30% of jobs re-query a resident matrix with a new target set): Poisson
    arrivals at `rate` jobs/s into a session of `slots` slots and a
    `budget_gb` HBM arena.  Reports p50/p90/p99 job latency (submit -> results
    in host memory), the same jobs submitted all at once (throughput at
    saturation), and the same jobs as plain one-shot C-ABI calls back to back."""
    from hpdex_amd import _capi, synthetic
    t0 = time.perf_counter()
    pool, jobs = synthetic.serving_workload(seed=seed, n_jobs=n_jobs)
    log(f"serving: generated {len(pool)} matrices and {len(jobs)} jobs in "
        f"{time.perf_counter() - t0:.1f}s")
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    # session jobs hand their matrices over with int32 indices, as scipy and
    # AnnData hold them; the plain one-shot calls take the ABI's int64 arrays
    for m in pool:
        m["indices32"] = m["indices"].astype(np.int32)
        m["indptr32"] = m["indptr"].astype(np.int32)
    resident = {k: _capi.DeviceMatrix(pool[k]["data"], pool[k]["indices"], pool[k]["indptr"],
                                      pool[k]["n_rows"])
                for k in sorted({j["matrix"] for j in jobs if j["resident"]})}
    tests = [pool[j["matrix"]]["n_cols"] * j["n_targets"] for j in jobs]

    def submit(sess, j):
        m = pool[j["matrix"]]
        if j["resident"]:
            return sess.submit(j["group_id"], j["n_targets"], opts, matrix=resident[j["matrix"]])
        return sess.submit(j["group_id"], j["n_targets"], opts, data=m["data"],
                           indices=m["indices32"], indptr=m["indptr32"], n_rows=m["n_rows"])

    sess = _capi.Session(budget_bytes=budget_gb << 30, slots=slots)
    for j in jobs[:2]:                       # warm-up (first-call costs), not measured
        sess.wait(submit(sess, j))
    # Poisson arrivals
    rng = np.random.default_rng(seed + 1)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, len(jobs)))
    tickets = []
    t0 = time.perf_counter()
    for j, a in zip(jobs, arrivals):
        d = t0 + a - time.perf_counter()
        if d > 0:
            time.sleep(d)
        tickets.append(submit(sess, j))
    lat = [sess.wait(t)[2]["latency_ms"] for t in tickets]
    t_arr = time.perf_counter() - t0
    peak = sess.peak()
    # saturation: every job queued at once
    t0 = time.perf_counter()
    tickets = [submit(sess, j) for j in jobs]
    sat_lat = [sess.wait(t)[2]["latency_ms"] for t in tickets]
    t_sat = time.perf_counter() - t0
    batching = sess.batches()
    sess.close()
    # plain: one blocking C-ABI call per job (upload, kernels, download)
    plain = []
    t0 = time.perf_counter()
    for j in jobs:
        m = pool[j["matrix"]]
        t1 = time.perf_counter()
        _capi.mwu_csc(m["data"], m["indices"], m["indptr"], j["group_id"], j["n_targets"], opts,
                      n_rows=m["n_rows"])
        plain.append((time.perf_counter() - t1) * 1e3)
    t_plain = time.perf_counter() - t0
    for m in resident.values():
        m.close()
    pct = lambda x, q: float(np.percentile(np.asarray(x), q))
    out = {
        "workload": f"{len(jobs)} jobs over {len(pool)} matrices (cells x genes, density): "
                    + ", ".join(f"{m['n_rows']}x{m['n_cols']} {m['density']:.3f}" for m in pool)
                    + f"; {sum(j['resident'] for j in jobs)} jobs re-query a resident matrix; "
                      "zero_handling=min, tie + continuity, asymptotic",
        "job_indices": "int32 CSC indices in session jobs (as scipy/AnnData hold them); "
                       "the plain one-shot calls take the ABI's int64 arrays",
        "arrival_rate_per_s": rate, "slots": slots, "hbm_budget_gb": budget_gb,
        "p50_ms": pct(lat, 50), "p90_ms": pct(lat, 90), "p99_ms": pct(lat, 99),
        "mean_ms": float(np.mean(lat)), "tests": int(sum(tests)),
        "tests_per_s_at_rate": sum(tests) / t_arr,
        "saturated_tests_per_s": sum(tests) / t_sat,
        "saturated_p50_ms": pct(sat_lat, 50), "saturated_p99_ms": pct(sat_lat, 99),
        "plain_tests_per_s": sum(tests) / t_plain,
        "plain_p50_ms": pct(plain, 50), "plain_p99_ms": pct(plain, 99),
        "saturated_over_plain": (sum(tests) / t_sat) / (sum(tests) / t_plain),
        "peak_jobs_in_flight": peak["jobs_in_flight"],
        "merged_launches": batching,
        "peak_arena_gb": peak["arena_bytes"] / 2 ** 30,
    }
    log(f"serving: p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms at {rate}/s; "
        f"saturated {out['saturated_tests_per_s']:.3e} tests/s vs plain "
        f"{out['plain_tests_per_s']:.3e} tests/s")
    return out


ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}


def opts_of(w):
    from hpdex_amd import _capi
    o = w["opts"]
    return _capi.make_opts(tie_correction=o["tie_correction"], use_continuity=o["use_continuity"],
                           zero_handling=ZH[o["zero_handling"]],
                           method=1 if o["method"] == "exact" else 2)


def describe(name, w):
    o = w["opts"]
    return (f"{name}: {w['n_rows']} cells x {w['n_cols']} genes, nnz {int(w['indptr'][-1])}, "
            f"reference {int((w['group_id'] == 0).sum())} cells, {w['n_targets']} targets, "
            f"zero_handling={o['zero_handling']}, method={o['method']}")


def first_job(w):
    """One job on a matrix that is not resident yet: upload (H2D of the int64
    ABI arrays, column classification, packed entries, row conversion),
    the plan (labels, routing, graph capture), one pass and the results back
    in host memory -- what a single job on a fresh matrix costs.  The
    process's one-time costs (code objects loaded, the first graph launched,
    staging buffers and copy threads created) are paid first by small jobs
    and are not part of it."""
    from hpdex_amd import _capi, synthetic
    opts = opts_of(w)
    ws = synthetic.small_case()
    ms = _capi.DeviceMatrix(ws["data"], ws["indices"], ws["indptr"], ws["n_rows"])
    ps = _capi.Plan(ms, ws["group_id"], ws["n_targets"], opts, form=1)   # graph launches
    ps.run(1)
    ps.results()
    ps.close()
    ms.close()
    _capi.mwu_csc(w["data"][:1 << 21], w["indices"][:1 << 21],
                  np.minimum(w["indptr"], 1 << 21), w["group_id"], w["n_targets"], opts,
                  n_rows=w["n_rows"])
    t0 = time.perf_counter()
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    t1 = time.perf_counter()
    p = _capi.Plan(m, w["group_id"], w["n_targets"], opts, form=1)
    t2 = time.perf_counter()
    p.run(1)
    U1, P = p.results()
    t3 = time.perf_counter()
    p.close()
    m.close()
    return {"first_job_ms": (t3 - t0) * 1e3, "matrix_upload_ms": (t1 - t0) * 1e3,
            "plan_ms": (t2 - t1) * 1e3, "pass_and_results_ms": (t3 - t2) * 1e3}


def measure(name, w, steps, warmup, plain=True, cpu=True, cpu_target_s=12.0, m=None):
    """The workload on a resident matrix: `steps` replays of the batched form
    (graph), the per-kernel times and the dominant kernel's roofline, the
    plain form on the same resident matrix (launched eagerly), the plain
    one-shot call from host buffers, and the CPU baseline sample."""
    from hpdex_amd import _capi, synthetic
    C, T, R = w["n_cols"], w["n_targets"], w["n_rows"]
    opts = opts_of(w)
    own = m is None
    if own:
        m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], R)
    # one job as it comes: every plan here sorts its own reference (the
    # matrix's kept reference tables are measured by requery_bench)
    keep = int(_capi.get_tuning().ref_cache)
    _capi.set_tuning(ref_cache=0)
    m.ref_cache_clear()
    pb = _capi.Plan(m, w["group_id"], T, opts, form=1)
    pb.run(warmup)
    log(f"{name}: timing {steps} steps (batched)")
    ms_total, graph_ms = pb.run(steps)
    ms_step = ms_total / steps
    U1b, Pb = pb.results()
    tests = C * T
    value = tests / (ms_step / 1e3)
    vbytes = 4 if w["data"].dtype == np.float32 else 8
    info = pb.info()
    # roofline of the dominant kernel (the largest share of the plain launch
    # sequence, HIP events between its kernels): its mean launch duration
    # from HIP events around launches of it alone, back to back on the
    # plan's stream
    stages = pb.stage_times(max(3, steps // 2))
    if stages["count"] >= max(stages["split"], stages["reference"], stages["target"]):
        dom = "count"
        B = synthetic.pass_bytes(info["count_nnz"], info["count_columns"], R, T, vbytes)
    else:
        dom = max(("split", "reference", "target"), key=lambda k: stages[k])
        B = synthetic.algorithmic_bytes(w, vbytes)
    dom_ms = pb.kernel_ms(dom, iters=max(10, steps))
    dom = dom + "_kernel"
    achieved = B / (dom_ms / 1e3) / 1e9
    step_B = synthetic.algorithmic_bytes(w, vbytes)
    log(f"{name}: {ms_step:.3f} ms/step, {value:.3e} tests/s; kernels {stages}; "
        f"dominant {dom} {achieved:.0f} GB/s algorithmic")
    out = {"workload": describe(name, w), "value": value, "unit": "tests/s",
           "ms_per_step": ms_step, "graph_ms": graph_ms, "tests_per_step": tests,
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": traffic_from_profiles(dom, name), "kernel": dom,
                        "algorithmic_bytes": B, "kernel_ms": dom_ms},
           "step_roofline_frac": step_B / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS,
           "kernel_ms": stages, "routing": info}
    if plain:
        pp = _capi.Plan(m, w["group_id"], T, opts, form=0)
        pp.run(2)
        pms, pker = pp.run(max(3, steps // 5))
        U1p, Pp = pp.results()
        same = bool(np.array_equal(U1p, U1b) and np.array_equal(Pp, Pb))
        pp.close()
        t0 = time.perf_counter()
        _capi.mwu_csc(w["data"], w["indices"], w["indptr"], w["group_id"], T, opts, n_rows=R)
        e2e = time.perf_counter() - t0
        out["plain"] = {"resident_ms": pker, "resident_tests_per_s": tests / (pker / 1e3),
                        "batched_over_plain_resident": value / (tests / (pker / 1e3)),
                        "per_call_ms": e2e * 1e3, "per_call_tests_per_s": tests / e2e,
                        "per_call_note": "one-shot C-ABI call from host buffers: upload of "
                                         "the int64 ABI arrays over PCIe, kernels, download",
                        "bit_identical_to_batched": same}
        log(f"{name}: plain resident {pker:.3f} ms (identical {same}); "
            f"one-shot call {e2e * 1e3:.1f} ms")
    pb.close()
    _capi.set_tuning(ref_cache=keep)
    if info["split_columns"] and keep:
        out["requery"] = requery_bench(name, w, m, opts, steps)
    if own:
        m.close()
    if cpu:
        log(f"{name}: CPU baseline (oracle)")
        out["cpu_baseline"] = cpu_baseline(w, target_s=cpu_target_s)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        log(f"{name}: CPU {out['cpu_baseline']['value']:.3e} tests/s")
    return out


def requery_bench(name, w, m, opts, steps, seed=11):
    """A new target set against the same reference on a resident matrix (the
    reference rows fixed, the targets' members redrawn with the same sizes):
    after the first job, the matrix keeps each sort-path column's sorted
    reference, and the follow-up job skips the reference sort.  Times the
    follow-up in the batched form with the kept tables and with them turned
    off, and checks that both give the same results bit for bit."""
    from hpdex_amd import _capi
    T = w["n_targets"]
    gid = w["group_id"].copy()
    tgt = np.nonzero(gid > 0)[0]
    gid[tgt] = w["group_id"][tgt][np.random.default_rng(seed).permutation(len(tgt))]
    m.ref_cache_clear()
    first = _capi.Plan(m, w["group_id"], T, opts, form=1)
    first.run(1)   # the matrix keeps this job's sorted reference
    first.close()
    kept = m.ref_cache_info()
    timed = {}
    for use in (1, 0):
        _capi.set_tuning(ref_cache=use)
        p = _capi.Plan(m, gid, T, opts, form=1)
        cached_cols = p.info()["ref_cached_columns"]
        p.run(2)
        ms_total, _ = p.run(steps)
        stages = p.stage_times(3)
        timed[use] = (ms_total / steps, p.results(), cached_cols, stages)
        p.close()
    _capi.set_tuning(ref_cache=1)
    same = bool(np.array_equal(timed[1][1][0], timed[0][1][0]) and
                np.array_equal(timed[1][1][1], timed[0][1][1]))
    log(f"{name}: re-query with the kept reference {timed[1][0]:.3f} ms/step "
        f"({timed[1][2]} columns, {kept['bytes'] / 2**20:.0f} MiB kept), without "
        f"{timed[0][0]:.3f} ms/step; identical {same}")
    return {"workload": "new target set (same sizes, members redrawn) against the same "
                        "reference rows on the resident matrix",
            "ms_per_step_kept_reference": timed[1][0],
            "ms_per_step_sorted_again": timed[0][0],
            "speedup": timed[0][0] / timed[1][0],
            "reference_kernel_ms": {"kept": timed[1][3]["reference"],
                                    "sorted_again": timed[0][3]["reference"]},
            "ref_cached_columns": int(timed[1][2]), "kept_columns": kept["columns"],
            "kept_bytes": kept["bytes"], "bit_identical": same}


def batch_bench(seed=1005, n_jobs=64, reps=5):
    """Many jobs on resident matrices, batched against plain: the serving
    workload's jobs (4 matrices of the serving distribution, each job a new
    target set of 5-500 targets) with every matrix in HBM.  Plain: one plan
    per job in the plain form, run one after another (each run launches its
    kernels eagerly and waits).  Batched: the same plans as one hpdex_batch
    (one queue of every job's count-kernel items, largest first, one graph).
    Wall time of `reps` passes over all jobs; results bit-identical."""
    from hpdex_amd import _capi, synthetic
    pool, jobs = synthetic.serving_workload(seed=seed, n_jobs=n_jobs)
    opts = _capi.make_opts(tie_correction=True, use_continuity=True, zero_handling=1)
    mats = [_capi.DeviceMatrix(m["data"], m["indices"], m["indptr"], m["n_rows"]) for m in pool]
    plans = [_capi.Plan(mats[j["matrix"]], j["group_id"], j["n_targets"], opts, form=0)
             for j in jobs]
    tests = sum(pool[j["matrix"]]["n_cols"] * j["n_targets"] for j in jobs)
    for p in plans:                      # warm-up
        p.run(1)
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in plans:
            p.run(1)
    t_plain = (time.perf_counter() - t0) / reps
    want = [p.results() for p in plans]
    b = _capi.Batch(plans)
    b.run(2)
    t0 = time.perf_counter()
    b.run(reps)
    t_batch = (time.perf_counter() - t0) / reps
    same = all(np.array_equal(p.results()[0], w[0]) and np.array_equal(p.results()[1], w[1])
               for p, w in zip(plans, want))
    info = b.info()
    b.close()
    for p in plans:
        p.close()
    for m in mats:
        m.close()
    out = {"workload": f"{len(jobs)} jobs of the serving distribution on {len(pool)} resident "
                       "matrices (" + ", ".join(f"{m['n_rows']}x{m['n_cols']}" for m in pool) +
                       "), zero_handling=min, tie + continuity, asymptotic",
           "tests": int(tests), "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3,
           "plain_tests_per_s": tests / t_plain, "batched_tests_per_s": tests / t_batch,
           "batched_over_plain": t_plain / t_batch, "bit_identical": bool(same),
           "count_items": info["count_items"]}
    log(f"batch: {len(jobs)} jobs, plain {t_plain * 1e3:.2f} ms, batched {t_batch * 1e3:.2f} ms "
        f"({out['batched_over_plain']:.2f}x), identical {same}")
    return out


def batch_target_sets(w, n_sets, reps=5, seed=7):
    """Several target sets against one resident matrix (the reference fixed,
    the targets redrawn per set with the same sizes), batched against plain
    as in batch_bench."""
    from hpdex_amd import _capi
    rng = np.random.default_rng(seed)
    opts = opts_of(w)
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], w["n_rows"])
    gid0 = w["group_id"]
    plans = []
    for _ in range(n_sets):
        gid = gid0.copy()
        tgt = np.nonzero(gid0 > 0)[0]
        gid[tgt] = gid0[tgt][rng.permutation(len(tgt))]   # same sizes, new members
        plans.append(_capi.Plan(m, gid, w["n_targets"], opts, form=0))
    tests = n_sets * w["n_cols"] * w["n_targets"]
    for p in plans:
        p.run(1)
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in plans:
            p.run(1)
    t_plain = (time.perf_counter() - t0) / reps
    want = [p.results() for p in plans]
    b = _capi.Batch(plans)
    b.run(2)
    t0 = time.perf_counter()
    b.run(reps)
    t_batch = (time.perf_counter() - t0) / reps
    same = all(np.array_equal(p.results()[0], x[0]) and np.array_equal(p.results()[1], x[1])
               for p, x in zip(plans, want))
    b.close()
    for p in plans:
        p.close()
    m.close()
    log(f"batch of {n_sets} target sets: plain {t_plain * 1e3:.2f} ms, batched "
        f"{t_batch * 1e3:.2f} ms ({t_plain / t_batch:.2f}x), identical {same}")
    return {"workload": f"{n_sets} target sets against one resident matrix: {describe('', w)}",
            "tests": int(tests), "plain_ms": t_plain * 1e3, "batched_ms": t_batch * 1e3,
            "plain_tests_per_s": tests / t_plain, "batched_tests_per_s": tests / t_batch,
            "batched_over_plain": t_plain / t_batch, "bit_identical": bool(same)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="histogram",
                    help="histogram (default), small, float, float-mix, exact")
    ap.add_argument("--workload-cache", default=None,
                    help="directory to keep the generated workload's arrays in (.npy) and read "
                         "them from on later runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the float (min, mix) and exact workloads reported beside the "
                         "headline")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the many-jobs batched-vs-plain measurement")
    ap.add_argument("--stream-limit-gb", type=float, default=0.0,
                    help="also stream the workload's DE from memory-mapped files under this "
                         "memory limit (GB) and compare with the one-shot call")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip the host-vs-device ingest measurement")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the serving measurement (p50/p99 job latency)")
    ap.add_argument("--serving-jobs", type=int, default=64)
    ap.add_argument("--serving-rate", type=float, default=8.0)
    ap.add_argument("--count-lds", type=int, default=None,
                    help="tuning: LDS bytes of the count kernel (0 sends every column to the "
                         "split/reference/target kernels)")
    ap.add_argument("--count-block", type=int, default=None,
                    help="tuning: workgroup size of the count kernel (256, 512, 1024)")
    ap.add_argument("--autotune", action="store_true",
                    help="time every candidate tuning on this workload before the timed steps "
                         "and run them under the fastest (the default, unless --no-autotune, "
                         "--count-lds, --count-block or --tuning is given)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run the timed steps under the built-in tuning")
    ap.add_argument("--tuning", default=None,
                    help="apply a tuning JSON written by hpdex_amd.autotune.save (with "
                         "--autotune: write the chosen tuning there)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run the streaming copy kernel over this many MiB (3 launches "
                         "after one warm launch) and print its GB/s: the calibration run for "
                         "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(genes would shard by column; see the design notes).", file=sys.stderr)
        sys.exit(2)

    from hpdex_amd import _capi
    if args.copy_probe_mib:
        gbs = _capi.bandwidth_probe(args.copy_probe_mib << 20, iters=3)
        print(json.dumps({"copy_probe_bytes": args.copy_probe_mib << 20, "gbs": gbs}), flush=True)
        return
    if args.count_lds is not None or args.count_block is not None:
        _capi.set_tuning(count_lds_bytes=args.count_lds, count_block=args.count_block)
    log(f"generating workload {args.workload}")
    t0 = time.perf_counter()
    w = load_workload(args.workload, args.workload_cache)
    C, T, R = w["n_cols"], w["n_targets"], w["n_rows"]
    nnz = int(w["indptr"][-1])
    log(f"generated in {time.perf_counter() - t0:.1f}s: {R} cells x {C} genes, nnz={nnz}, T={T}")
    log(f"device {_capi.device_info()}")
    extra = {}
    if not args.no_plain:
        extra["first_job"] = first_job(w)
        log(f"first job on a fresh matrix: {extra['first_job']}")
    m = _capi.DeviceMatrix(w["data"], w["indices"], w["indptr"], R)
    log(f"resident matrix: {m.device_bytes() / 1e9:.2f} GB")
    tuned = None
    manual = args.count_lds is not None or args.count_block is not None or args.tuning
    if args.autotune or not (args.no_autotune or manual):
        from hpdex_amd import autotune
        t0 = time.perf_counter()
        # knobs given by hand are kept: the sweep covers the others only
        space = {k: v for k, v in autotune.space_for(m).items()
                 if not (k == "count_lds_bytes" and args.count_lds is not None)
                 and not (k == "count_block" and args.count_block is not None)}
        try:
            res = autotune.autotune(m, w["group_id"], T, opts_of(w), space=space, log=log)
        except Exception as e:       # the tuning in force was restored: time under it
            tuned = {"error": str(e), "tuning": autotune.current()}
            log(f"autotune failed, timing the tuning in force: {e}")
        else:
            n_ok = sum(t["ms"] is not None for t in res["trials"])
            tuned = {"tuning": res["best"], "ms": res["ms"], "candidates": len(res["trials"]),
                     "runnable": n_ok, "seconds": time.perf_counter() - t0}
            log(f"autotune: {tuned}")
            if args.tuning:
                autotune.save(res, args.tuning)
    elif args.tuning:
        from hpdex_amd import autotune
        tuned = {"tuning": autotune.load(args.tuning), "from": args.tuning}
    head = measure(args.workload, w, args.steps, args.warmup, plain=not args.no_plain,
                   cpu=not args.no_cpu_baseline, m=m)
    m.close()
    if not args.no_ingest and args.workload in ("histogram", "small"):
        log("ingest: CSR with int32 indices, host route vs device route")
        extra["ingest"] = ingest_bench(w)
        ig = extra["ingest"]
        log(f"ingest: host {ig['host_route_ms']:.0f} ms (tocsc + int64 {ig['host_tocsc_and_int64_ms']:.0f}, "
            f"upload {ig['host_csc_upload_ms']:.0f}), device {ig['device_route_ms']:.0f} ms")
    if args.stream_limit_gb > 0:
        log(f"streaming DE under {args.stream_limit_gb} GB")
        extra["stream"] = stream_bench(w, args.stream_limit_gb)
        sb = extra["stream"]
        log(f"stream: {sb['runs']} runs, {sb['stream_ms']:.0f} ms vs one-shot {sb['one_shot_ms']:.0f} ms, "
            f"identical {sb['identical']}")
    del w
    if not args.no_configs and args.workload == "histogram":
        extra["configs"] = {}
        for name in ("float", "float-mix", "exact"):
            log(f"workload {name}")
            wx = load_workload(name, args.workload_cache)
            extra["configs"][name] = measure(name, wx, max(3, args.steps // 5), 2,
                                             plain=not args.no_plain,
                                             cpu=not args.no_cpu_baseline, cpu_target_s=6.0)
            del wx
    if not args.no_batch:
        from hpdex_amd import synthetic
        log("many jobs on resident matrices: batched vs plain")
        extra["batch"] = {"serving_mix": batch_bench(n_jobs=args.serving_jobs),
                          "small_target_sets": batch_target_sets(synthetic.small_case(), 64)}
        if args.workload == "histogram":
            wh = load_workload("histogram", args.workload_cache)
            extra["batch"]["histogram_target_sets"] = batch_target_sets(wh, 8)
            del wh
    serving = None
    if not args.no_serving:
        log("serving workload")
        serving = serving_bench(n_jobs=args.serving_jobs, rate=args.serving_rate)
    line = {
        "metric": "gene\u00d7target MWU tests/sec on 1 MI355X, batched vs plain; p50/p99 job latency",
        "value": head["value"],
        "unit": "tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-keys/int64-U/f64-p",
        "data": "synthetic (seeded generator, hpdex_amd/synthetic.py)",
        "config": {"workload": head["workload"], "cells": R, "genes": C, "targets": T,
                   "nnz": nnz, "tests_per_step": head["tests_per_step"],
                   "parallelism": "single GPU"},
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "batched_over_plain_resident": head.get("plain", {}).get("batched_over_plain_resident"),
        "kernel_ms": head["kernel_ms"],
        "routing": head["routing"],
        "graph_ms": head["graph_ms"],
        "step_roofline_frac": head["step_roofline_frac"],
        "plain": head.get("plain"),
        "serving": serving,
    }
    if tuned is not None:
        line["autotune"] = tuned
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

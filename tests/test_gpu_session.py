"""Serving session on the GPU: jobs of different sizes, paths and modes run
concurrently (several slots, one HBM arena) give exactly the results of the
same jobs run alone through the one-shot C-ABI call; admission respects the
arena budget; re-queries of a resident matrix with new target sets match."""
import sys

import numpy as np
import pytest

import oracle
from fixtures import rel_close
from hpdex_amd import _capi, synthetic

pytestmark = pytest.mark.gpu

ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _matrix(seed, R, C, dens, values):
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, dens))
    return rng, dict(data=values(rng, len(rows)).astype(np.float32), indices=rows,
                     indptr=indptr, n_rows=R, n_cols=C)


def _groups(rng, R, T, ref_frac=0.2, ignored=0.1):
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    nref, nign = int(R * ref_frac), int(R * ignored)
    gid[perm[:nref]] = 0
    gid[perm[nref + nign:]] = rng.integers(1, T + 1, R - nref - nign)
    return gid


def _jobs():
    """(name, matrix dict, group_id, T, opts, resident?) of a varied job mix."""
    jobs = []
    rng, counts = _matrix(31, 8000, 64, 0.15, lambda r, n: r.negative_binomial(0.8, 0.15, n) + 1)
    jobs.append(("count-min", counts, _groups(rng, 8000, 20), 20,
                 _capi.make_opts(True, True, ZH["min"]), False))
    rng, floats = _matrix(32, 6000, 40, 0.3, lambda r, n: np.round(r.normal(0, 2, n), 2))
    jobs.append(("float-mix", floats, _groups(rng, 6000, 8), 8,
                 _capi.make_opts(True, False, ZH["mix"]), False))
    rng, zeros = _matrix(33, 5000, 30, 0.3, lambda r, n: r.integers(0, 5, n))
    jobs.append(("count-none", zeros, _groups(rng, 5000, 6), 6,
                 _capi.make_opts(True, True, ZH["none"]), False))
    ex = synthetic.exact_case(n_cells=2000, n_genes=24, n_ref=40, n_targets=30)
    jobs.append(("exact-min", dict(data=ex["data"], indices=ex["indices"], indptr=ex["indptr"],
                                   n_rows=ex["n_rows"], n_cols=ex["n_cols"]),
                 ex["group_id"], ex["n_targets"],
                 _capi.make_opts(True, False, ZH["min"], method=1), False))
    # two target sets against one resident matrix
    rng, res = _matrix(34, 12000, 48, 0.1, lambda r, n: r.negative_binomial(1.0, 0.2, n) + 1)
    for k, T in enumerate((12, 30)):
        jobs.append((f"resident-{k}", res, _groups(np.random.default_rng(40 + k), 12000, T), T,
                     _capi.make_opts(True, True, ZH["max"]), True))
    return jobs


def _alone(m, gid, T, opts):
    return _capi.mwu_csc(m["data"], m["indices"], m["indptr"], gid, T, opts, n_rows=m["n_rows"])


def _submit(sess, job, resident):
    name, m, gid, T, opts, is_res = job
    if is_res:
        key = id(m)
        if key not in resident:
            resident[key] = _capi.DeviceMatrix(m["data"], m["indices"], m["indptr"], m["n_rows"])
        return sess.submit(gid, T, opts, matrix=resident[key])
    return sess.submit(gid, T, opts, data=m["data"], indices=m["indices"], indptr=m["indptr"],
                       n_rows=m["n_rows"])


def test_session_mix_matches_jobs_alone():
    jobs = _jobs()
    say("session: jobs alone")
    alone = [_alone(m, gid, T, opts) for _, m, gid, T, opts, _ in jobs]
    say("session: jobs in the mix (4 slots, each job twice)")
    sess = _capi.Session(budget_bytes=4 << 30, slots=4)
    resident = {}
    tickets = [(i, _submit(sess, jobs[i], resident)) for i in list(range(len(jobs))) * 2]
    for i, t in tickets:
        U1, P, times = sess.wait(t)
        assert np.array_equal(U1, alone[i][0]), jobs[i][0]
        assert np.array_equal(P, alone[i][1]), jobs[i][0]
        assert times["latency_ms"] >= times["run_ms"] >= 0.0
    peak = sess.peak()
    assert 1 <= peak["jobs_in_flight"] <= 4
    assert peak["arena_bytes"] <= 4 << 30
    # every job's kernels ran in one of the launcher's merged launches
    bt = sess.batches()
    assert bt["jobs"] == len(tickets) and 1 <= bt["launches"] <= len(tickets)
    assert 1 <= bt["max_jobs"] <= 4
    sess.close()
    for m in resident.values():
        m.close()
    # one job of the mix against the oracle (U1 bit-exact, p to 1e-12)
    name, m, gid, T, opts, _ = jobs[1]
    o = oracle.mwu(m["data"], m["indices"], m["indptr"], gid, T, n_rows=m["n_rows"],
                   tie_correction=True, use_continuity=False, zero_handling="mix")
    assert np.array_equal(alone[1][0], o[0]) and rel_close(alone[1][1], o[1], rtol=1e-12)


def test_session_admission_respects_budget():
    """A budget a little above the largest job's footprint still runs every job
    (one or two at a time) with the same results; a job above the budget is
    refused when submitted."""
    jobs = _jobs()[:3]
    say("session: footprints")
    sess = _capi.Session(budget_bytes=2 << 30, slots=2)
    feet = []
    for job in jobs:
        t = _submit(sess, job, {})
        feet.append(sess.wait(t)[2]["footprint_bytes"])
    sess.close()
    budget = max(feet) + (1 << 20)
    say(f"session: budget {budget} bytes, footprints {feet}")
    sess = _capi.Session(budget_bytes=budget, slots=3)
    tickets = [_submit(sess, job, {}) for job in jobs * 2]
    for k, t in enumerate(tickets):
        U1, P, _ = sess.wait(t)
        _, m, gid, T, opts, _ = jobs[k % len(jobs)]
        a = _alone(m, gid, T, opts)
        assert np.array_equal(U1, a[0]) and np.array_equal(P, a[1])
    assert sess.peak()["arena_bytes"] <= budget
    sess.close()
    small = _capi.Session(budget_bytes=2 << 20, slots=1)
    with pytest.raises(_capi.HpdexError, match="budget"):
        _submit(small, jobs[0], {})
    small.close()


def test_session_jobs_with_int32_indices_and_csr():
    """Jobs that bring their matrix with int32 indices, or as CSR, return the
    results of the same matrix as an int64 CSC."""
    import scipy.sparse as sp
    rng, m = _matrix(36, 5000, 40, 0.15, lambda r, n: r.negative_binomial(0.8, 0.2, n) + 1)
    m["data"][::7] += 0.5                       # some float columns: both paths
    gid = _groups(rng, 5000, 6)
    opts = _capi.make_opts(True, True, ZH["mix"])
    a = _alone(m, gid, 6, opts)
    csc = sp.csc_matrix((m["data"], m["indices"], m["indptr"]), shape=(5000, 40))
    csr = csc.tocsr()
    say("session: int32 CSC and CSR jobs")
    sess = _capi.Session(budget_bytes=1 << 30, slots=2)
    t1 = sess.submit(gid, 6, opts, data=m["data"], indices=m["indices"].astype(np.int32),
                     indptr=m["indptr"].astype(np.int32), n_rows=5000)
    t2 = sess.submit(gid, 6, opts, data=csr.data, indices=csr.indices.astype(np.int32),
                     indptr=csr.indptr.astype(np.int32), n_rows=5000, n_cols=40,
                     layout=_capi.LAYOUT_CSR)
    for t in (t1, t2):
        U1, P, _ = sess.wait(t)
        assert np.array_equal(U1, a[0]) and np.array_equal(P, a[1])
    sess.close()


def test_session_reports_job_errors():
    """A job whose reference has fewer than two cells fails with the
    reference's message; the session keeps serving."""
    rng, m = _matrix(35, 3000, 10, 0.2, lambda r, n: r.integers(1, 9, n))
    gid = _groups(rng, 3000, 4)
    bad = gid.copy()
    bad[bad == 0] = -1
    bad[0] = 0
    opts = _capi.make_opts(True, True, ZH["none"])
    sess = _capi.Session(budget_bytes=1 << 30, slots=2)
    t_bad = _submit(sess, ("bad", m, bad, 4, opts, False), {})
    t_ok = _submit(sess, ("ok", m, gid, 4, opts, False), {})
    with pytest.raises(_capi.HpdexError, match="Sample too small for reference"):
        sess.wait(t_bad)
    U1, P, _ = sess.wait(t_ok)
    a = _alone(m, gid, 4, opts)
    assert np.array_equal(U1, a[0]) and np.array_equal(P, a[1])
    sess.close()


def test_serving_workload_distribution():
    """Jobs drawn from the serving distribution (cells U[5k, 200k], genes
    U[2k, 20k], density U[3 %, 12 %], NB(LogNormal(0, 1), r = 0.7) + 1 counts,
    a reference of 10 % of the cells, 5-500 targets, 30 % re-querying a
    resident matrix) submitted at once into a 4-slot session: each job equals
    the oracle (U1 bit for bit, p within 1e-12) on a sample of its genes that
    includes its densest ones, and a few jobs equal the same job run alone."""
    import os
    say("serving workload generate")
    pool, jobs = synthetic.serving_workload(seed=1105, n_matrices=3, n_jobs=12)
    opts = _capi.make_opts(True, True, ZH["min"])
    resident = {k: _capi.DeviceMatrix(pool[k]["data"], pool[k]["indices"], pool[k]["indptr"],
                                      pool[k]["n_rows"])
                for k in sorted({j["matrix"] for j in jobs if j["resident"]})}
    assert resident, "the sample has resident re-queries"
    sess = _capi.Session(budget_bytes=48 << 30, slots=4)
    tickets = []
    for j in jobs:
        m = pool[j["matrix"]]
        if j["resident"]:
            tickets.append(sess.submit(j["group_id"], j["n_targets"], opts,
                                       matrix=resident[j["matrix"]]))
        else:
            tickets.append(sess.submit(j["group_id"], j["n_targets"], opts, data=m["data"],
                                       indices=m["indices"].astype(np.int32),
                                       indptr=m["indptr"].astype(np.int32), n_rows=m["n_rows"]))
    say(f"serving workload: {len(jobs)} jobs submitted")
    results = [sess.wait(t)[:2] for t in tickets]
    assert sess.peak()["jobs_in_flight"] >= 2
    bt = sess.batches()
    assert bt["jobs"] == len(jobs) and bt["launches"] <= len(jobs)
    say(f"serving workload: {bt}")
    sess.close()
    for m in resident.values():
        m.close()
    threads = min(16, os.cpu_count() or 1)
    for k, (j, (U1, P)) in enumerate(zip(jobs, results)):
        m = pool[j["matrix"]]
        T, C = j["n_targets"], m["n_cols"]
        nnz = np.diff(m["indptr"])
        cols = np.unique(np.concatenate([np.argsort(nnz)[-5:], np.linspace(0, C - 1, 25)
                                         .astype(np.int64)]))
        sub = synthetic.subset_columns(dict(m, group_id=j["group_id"], n_targets=T), cols)
        say(f"serving job {k}: {m['n_rows']} x {C}, T={T}, oracle on {len(cols)} genes")
        oU1, oP = oracle.mwu(sub["data"], sub["indices"], sub["indptr"], j["group_id"], T,
                             n_rows=m["n_rows"], tie_correction=True, use_continuity=True,
                             zero_handling="min", threads=threads)
        gU1 = U1.reshape(C, T)[cols].reshape(-1)
        gP = P.reshape(C, T)[cols].reshape(-1)
        assert np.array_equal(gU1, oU1), k
        assert rel_close(gP, oP, rtol=1e-12), k
        if k < 3:
            a = _alone(m, j["group_id"], T, opts)
            assert np.array_equal(U1, a[0]) and np.array_equal(P, a[1]), k

"""Batches of plans on the GPU: several jobs -- different target sets on one
resident matrix, jobs on other matrices, float and double values, count path
and sort path, exact p -- run as one launch sequence give every plan exactly
the results it gives alone, on every replay, and equal to the oracle."""
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


def _matrix(seed, R, C, dens, values, dtype=np.float32):
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, dens))
    return rng, dict(data=values(rng, len(rows)).astype(dtype), indices=rows, indptr=indptr,
                     n_rows=R, n_cols=C)


def _groups(rng, R, T, ref_frac=0.2):
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    nref = int(R * ref_frac)
    gid[perm[:nref]] = 0
    gid[perm[nref:]] = rng.integers(1, T + 1, R - nref)
    return gid


def _jobs():
    """(matrix key, group_id, T, opts) of a mix, and the matrices."""
    mats = {}
    rng, mats["counts"] = _matrix(41, 9000, 120, 0.12,
                                  lambda r, n: r.negative_binomial(0.8, 0.15, n) + 1)
    jobs = [("counts", _groups(np.random.default_rng(50 + k), 9000, T), T,
             _capi.make_opts(True, True, ZH[mode]))
            for k, (T, mode) in enumerate(((12, "min"), (40, "mix"), (3, "none"), (200, "max")))]
    rng, mats["floats"] = _matrix(42, 7000, 60, 0.3, lambda r, n: np.round(r.normal(0, 2, n), 2))
    jobs.append(("floats", _groups(rng, 7000, 25), 25, _capi.make_opts(True, False, ZH["mix"])))
    rng, mats["doubles"] = _matrix(43, 5000, 30, 0.4, lambda r, n: r.integers(-20, 60, n),
                                   dtype=np.int32)
    jobs.append(("doubles", _groups(rng, 5000, 6), 6, _capi.make_opts(True, True, ZH["min"])))
    ex = synthetic.exact_case(n_cells=2000, n_genes=24, n_ref=40, n_targets=30)
    mats["exact"] = dict(data=ex["data"], indices=ex["indices"], indptr=ex["indptr"],
                         n_rows=ex["n_rows"], n_cols=ex["n_cols"])
    jobs.append(("exact", ex["group_id"], ex["n_targets"],
                 _capi.make_opts(True, False, ZH["min"], method=1)))
    return mats, jobs


def test_batch_equals_plans_alone():
    mats, jobs = _jobs()
    dm = {k: _capi.DeviceMatrix(m["data"], m["indices"], m["indptr"], m["n_rows"])
          for k, m in mats.items()}
    say("batch: plans alone")
    plans, alone = [], []
    for key, gid, T, opts in jobs:
        p = _capi.Plan(dm[key], gid, T, opts, form=1)
        p.run(1)
        alone.append(p.results())
        plans.append(p)
    say("batch: together")
    b = _capi.Batch(plans)
    info = b.info()
    assert info["count_items"] > 0 and info["split_columns"] > 0
    for rep in range(3):
        # the plans' outputs are overwritten by each replay
        for p in plans:
            p.run(1)
        ms = b.run(2)
        assert ms > 0
        for k, p in enumerate(plans):
            U1, P = p.results()
            assert np.array_equal(U1, alone[k][0]), (rep, jobs[k][0])
            assert np.array_equal(P, alone[k][1]), (rep, jobs[k][0])
    b.close()
    # against the oracle
    for (key, gid, T, opts), (U1, P) in zip(jobs, alone):
        m = mats[key]
        zh = {v: k for k, v in ZH.items()}[opts.zero_handling]
        o = oracle.mwu(m["data"], m["indices"], m["indptr"], gid, T, n_rows=m["n_rows"],
                       tie_correction=bool(opts.tie_correction),
                       use_continuity=bool(opts.use_continuity), zero_handling=zh,
                       method="exact" if opts.method == 1 else "asymptotic")
        assert np.array_equal(U1, o[0]), key
        assert rel_close(P, o[1], rtol=1e-12), key
    for p in plans:
        p.close()
    for m in dm.values():
        m.close()


def test_batch_rejects_bad_plans():
    rng, m = _matrix(44, 3000, 10, 0.3, lambda r, n: r.integers(1, 9, n))
    dm = _capi.DeviceMatrix(m["data"], m["indices"], m["indptr"], 3000)
    gid = _groups(rng, 3000, 4)
    p = _capi.Plan(dm, gid, 4, _capi.make_opts(True, True, ZH["min"]))
    with pytest.raises(_capi.HpdexError, match="twice"):
        _capi.Batch([p, p])
    q = _capi.Plan(dm, gid, 4, _capi.make_opts(True, False, ZH["none"], method=1))
    with pytest.raises(_capi.HpdexError, match="zero_handling none"):
        _capi.Batch([p, q])
    with pytest.raises(_capi.HpdexError, match="at least one plan"):
        _capi.Batch([])
    p.close(); q.close(); dm.close()

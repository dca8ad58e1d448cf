"""The gfx950 kernel source executed on the host (one thread per lane; see
hpdex_amd/csrc/cpu_sim) against the oracle: U1, tie term and sample sizes
bit-identical, p identical.  CPU only; covers the histogram and sort paths,
every zero mode, reference slices above one register run, target slices
split into several register runs, and 8-byte keys."""
import numpy as np
import pytest

import oracle
from fixtures import mwu_cases
from hpdex_amd import synthetic
from hpdex_amd.csrc.cpu_sim import sim

ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}
ALT = {"less": 0, "greater": 1, "two-sided": 2}


def run_both(data, indices, indptr, gid, T, R, mode, alt="two-sided", tie=True, cc=True,
             f64=False):
    U1, P, TIE, n1, n2, err, status = sim.mwu(data, indices, indptr, gid, T, R, mode=ZH[mode],
                                              alt=ALT[alt], tie=tie, cc=cc, f64=f64,
                                              timeout_s=120)
    assert err == 2 ** 64 - 1 and status == 0
    oU1, oP, oT, on1, on2 = oracle.mwu(data, indices, indptr, gid, T, n_rows=R,
                                       tie_correction=tie, use_continuity=cc,
                                       zero_handling=mode, alternative=alt, with_stats=True)
    assert np.array_equal(U1, oU1)
    assert np.array_equal(TIE, oT)
    assert np.array_equal(n1, on1) and np.array_equal(n2, on2)
    assert np.array_equal(P, oP)


SOME = [c for c in mwu_cases() if c["method"] == "asymptotic"][::5]


@pytest.mark.parametrize("case", SOME, ids=[c["name"] for c in SOME])
def test_sim_golden(case):
    run_both(case["data"], case["indices"], case["indptr"], case["group_id"], case["n_targets"],
             case["n_rows"], case["zero_handling"], case["alternative"], case["tie_correction"],
             case["use_continuity"])


def _rw(seed, R, C, T, dens, fn, ref_frac=0.3):
    rng = np.random.default_rng(seed)
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, dens))
    data = fn(rng, len(rows)).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    nref = int(R * ref_frac)
    gid[perm[:nref]] = 0
    gid[perm[nref:]] = rng.integers(1, T + 1, R - nref)
    return data, rows, indptr, gid


@pytest.mark.parametrize("mode", ["min", "none"])
def test_sim_sort_path_long_slices(mode):
    # reference ~900 keys (two register runs + an LDS merge stage),
    # targets ~1,000 keys (two register runs with cross-run tie counts)
    d, r, ip, g = _rw(5, 6000, 3, 2, 0.5, lambda rg, n: np.round(rg.normal(0, 2, n), 2))
    run_both(d, r, ip, g, 2, 6000, mode)


def test_sim_eight_byte_keys():
    d, r, ip, g = _rw(9, 1500, 3, 3, 0.5, lambda rg, n: np.round(rg.normal(0, 2, n), 1))
    run_both(d.astype(np.float64), r, ip, g, 3, 1500, "mix", f64=True)


def test_sim_histogram_path():
    d, r, ip, g = _rw(6, 4000, 4, 5, 0.3, lambda rg, n: rg.negative_binomial(0.8, 0.2, n) - 3)
    run_both(d, r, ip, g, 5, 4000, "mix")

"""The gfx950 kernel source executed on the host (one thread per lane; see
tests/cpu_sim) against the oracle: U1, tie term and sample sizes
bit-identical, p identical.  CPU only; covers the histogram and sort paths,
every zero mode, reference slices above one register run, target slices
split into several register runs, and 8-byte keys."""
import numpy as np
import pytest

import oracle
from fixtures import mwu_cases
from hpdex_amd import synthetic
from cpu_sim import sim

ZH = {"none": 0, "min": 1, "max": 2, "mix": 3}
ALT = {"less": 0, "greater": 1, "two-sided": 2}


def run_both(data, indices, indptr, gid, T, R, mode, alt="two-sided", tie=True, cc=True,
             f64=False, **kw):
    U1, P, TIE, n1, n2, err, status = sim.mwu(data, indices, indptr, gid, T, R, mode=ZH[mode],
                                              alt=ALT[alt], tie=tie, cc=cc, f64=f64,
                                              timeout_s=120, **kw)
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


@pytest.mark.parametrize("mode", ["min", "none", "mix"])
def test_sim_sort_path_many_small_targets(mode):
    # 150 targets of 2-40 cells (tiny or, in min/mix, empty slices, packed
    # several to a wave window) and four targets of ~400 cells (more than 64
    # keys: ranked alone), float values with ties and negatives; mode none
    # counts stored values only, so it gets a dense matrix (two per target)
    rng = np.random.default_rng(31)
    R, C, T = 8000, 6, 154
    dens = 0.97 if mode == "none" else 0.4
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, dens))
    data = np.round(rng.normal(0, 1.5, len(rows)), 1).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:600]] = 0
    sizes = np.concatenate([rng.integers(2 if mode != "none" else 6, 41, 150),
                            [400, 380, 420, 390]])
    at = 600
    for g, n in enumerate(sizes, start=1):
        gid[perm[at:at + n]] = g
        at += n
    run_both(data, rows, indptr, gid, T, R, mode)


@pytest.mark.parametrize("arrival", [True, False], ids=["groups-kept", "two-reads"])
@pytest.mark.parametrize("mode", ["min", "none"])
def test_sim_sort_path_sparse_columns(mode, arrival):
    # columns holding under a quarter of the rows are split through the
    # bit-packed labels: ~5,200 entries (two whole batches of 2,048 and a
    # partial one), ~2,000 (one partial batch) and, in mode min (mode none
    # needs two stored values per group), ~30 entries and an empty column;
    # rows in no group (label -1) interleaved
    rng = np.random.default_rng(33)
    R, T = 40000, 5
    dens = [0.13, 0.13, 0.05] + ([0.00075, 0.0] if mode == "min" else [])
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array(dens))
    data = np.round(rng.normal(0, 2, len(rows)), 2).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:9000]] = 0
    gid[perm[9000:30000]] = rng.integers(1, T + 1, 21000)
    run_both(data, rows, indptr, gid, T, R, mode, split_arrival=arrival)


def _dense(R, C):
    return np.tile(np.arange(R, dtype=np.int64), C), np.arange(C + 1, dtype=np.int64) * R


@pytest.mark.parametrize("mode", ["min", "none", "mix"])
def test_sim_sort_path_dense_slots(mode):
    # dense columns are split in one pass into slots sized by the groups' row
    # counts; explicit zeros and NaN among the stored values leave unused slot
    # entries after a group's keys, which the small-target windows step over
    rng = np.random.default_rng(34)
    R, C, T = 6000, 4, 200
    rows, indptr = _dense(R, C)
    data = np.round(rng.normal(0, 1.5, R * C), 1).astype(np.float32)
    z = rng.random(R * C)
    data[z < 0.1] = 0.0
    data[(z >= 0.1) & (z < 0.12)] = np.nan
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:500]] = 0
    at = 500
    for g, n in enumerate(rng.integers(4, 31, T), start=1):
        gid[perm[at:at + n]] = g
        at += n
    run_both(data, rows, indptr, gid, T, R, mode)


@pytest.mark.parametrize("mode", ["min", "none"])
def test_sim_sort_path_windows_with_a_large_reference(mode):
    # a reference of ~5,800 stored values: a window's sums may pass 2^32, so
    # the windows scan in 64 bits; 120 targets of 3-30 cells, ties
    rng = np.random.default_rng(40)
    R, C, T = 9000, 2, 120
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.full(C, 0.97))
    data = np.round(rng.normal(0, 1.0, len(rows)), 1).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:6000]] = 0
    at = 6000
    for g, n in enumerate(rng.integers(3, 31, T), start=1):
        gid[perm[at:at + n]] = g
        at += n
    run_both(data, rows, indptr, gid, T, R, mode)


@pytest.mark.parametrize("mode", ["min", "none"])
def test_sim_sort_path_dense_rows_out_of_order(mode):
    # three columns storing every row: in order (rows taken from positions),
    # reversed, and in order but for two rows swapped (both read their rows)
    rng = np.random.default_rng(39)
    R, T = 4000, 30
    rows = [np.arange(R), np.arange(R)[::-1].copy(), np.arange(R)]
    rows[2][[100, 2500]] = rows[2][[2500, 100]]
    rows = np.concatenate(rows).astype(np.int64)
    ip = np.arange(4, dtype=np.int64) * R
    data = np.round(rng.normal(0, 1.5, 3 * R), 1).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:600]] = 0
    gid[perm[600:2600]] = rng.integers(1, T + 1, 2000)
    run_both(data, rows, ip, gid, T, R, mode)


def test_sim_sort_path_dense_slots_beyond_lds():
    # 45,000 grouped rows: the slots (180 KB of keys) do not fit a workgroup's
    # LDS, so the one-pass split writes each key to its slot in HBM
    rng = np.random.default_rng(41)
    R, C, T = 46000, 2, 40
    rows, indptr = _dense(R, C)
    data = np.round(rng.normal(0, 1.5, R * C), 1).astype(np.float32)
    data[rng.random(R * C) < 0.05] = 0.0
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:6000]] = 0        # within the 8,192 reference keys kept on chip
    gid[perm[6000:45000]] = rng.integers(1, T + 1, 39000)
    run_both(data, rows, indptr, gid, T, R, "min")


def test_sim_sort_path_dense_slots_histogram_plan():
    # grouped values all small integers, one non-integer in a row of no group:
    # the column takes the sort path, whose split picks a histogram plan after
    # its keys are already in their slots
    rng = np.random.default_rng(35)
    R, C, T = 5000, 3, 40
    rows, indptr = _dense(R, C)
    data = rng.integers(-3, 12, R * C).astype(np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:800]] = 0
    gid[perm[800:3500]] = rng.integers(1, T + 1, 2700)
    for c in range(C):
        data[c * R + perm[4000 + c]] = 2.5
    U1, P, TIE, n1, n2, err, status = sim.mwu(data, rows, indptr, gid, T, R, mode=ZH["min"],
                                              timeout_s=120)
    assert sim.mwu.last_counted == 0
    assert err == 2 ** 64 - 1 and status == 0
    oU1, oP, oT, _, _ = oracle.mwu(data, rows, indptr, gid, T, n_rows=R, zero_handling="min",
                                   use_continuity=True, with_stats=True)
    assert np.array_equal(U1, oU1) and np.array_equal(TIE, oT) and np.array_equal(P, oP)


def test_sim_dense_slots_repeated_row_is_an_error():
    # a column storing one row twice (every row grouped, so the slots fill the
    # column's key range): the last group overflows its slot; the error is
    # reported and no key lands outside the column (checked by the sanitizer)
    rng = np.random.default_rng(36)
    R, T = 3000, 6
    gid = rng.integers(0, T + 1, R).astype(np.int32)
    last = np.nonzero(gid == T)[0]
    rows = np.arange(R, dtype=np.int64)
    other = np.nonzero(gid == 1)[0][0]
    rows[other] = last[0]
    rows.sort()
    data = rng.normal(0, 1, R).astype(np.float32)
    U1, P, TIE, n1, n2, err, status = sim.mwu(data, rows, np.array([0, R]), gid, T, R,
                                              mode=ZH["min"], timeout_s=60)
    assert err != 2 ** 64 - 1


def test_sim_eight_byte_keys():
    d, r, ip, g = _rw(9, 1500, 3, 3, 0.5, lambda rg, n: np.round(rg.normal(0, 2, n), 1))
    run_both(d.astype(np.float64), r, ip, g, 3, 1500, "mix", f64=True)


def test_sim_histogram_path():
    d, r, ip, g = _rw(6, 4000, 4, 5, 0.3, lambda rg, n: rg.negative_binomial(0.8, 0.2, n) - 3)
    run_both(d, r, ip, g, 5, 4000, "mix")


# ---- the count kernel (integer-valued columns in one pass) ----

def run_count(data, indices, indptr, gid, T, R, mode, count_lds, f64=False, hist_max=8192,
              expect_counted=None, block=512, expect_blocks=None, pack=True,
              expect_packed=None):
    U1, P, TIE, n1, n2, err, status = sim.mwu(data, indices, indptr, gid, T, R, mode=ZH[mode],
                                              count_lds=count_lds, f64=f64, hist_max=hist_max,
                                              block=block, timeout_s=120, pack=pack)
    if expect_packed is not None:
        assert sim.mwu.last_packed == expect_packed
    assert err == 2 ** 64 - 1 and status == 0
    oU1, oP, oT, on1, on2 = oracle.mwu(data, indices, indptr, gid, T, n_rows=R,
                                       zero_handling=mode, use_continuity=True, with_stats=True)
    assert np.array_equal(U1, oU1)
    assert np.array_equal(TIE, oT)
    assert np.array_equal(n1, on1) and np.array_equal(n2, on2)
    assert np.array_equal(P, oP)
    if expect_counted is not None:
        assert sim.mwu.last_counted == expect_counted
    if expect_blocks is not None:
        assert sim.mwu.last_blocks == expect_blocks


@pytest.mark.parametrize("pack", [True, False])
@pytest.mark.parametrize("block", [256, 512, 1024])
@pytest.mark.parametrize("mode", ["min", "max", "mix", "none"])
def test_sim_count_kernel_modes(mode, block, pack):
    # small signed integers with explicit zeros: one workgroup per column, read
    # as packed entries (one word per entry) or as rows and values
    d, r, ip, g = _rw(21, 3000, 5, 7, 0.4, lambda rg, n: rg.integers(-4, 9, n))
    run_count(d, r, ip, g, 7, 3000, mode, 76 * 1024, expect_counted=5, expect_blocks=5,
              block=block, pack=pack, expect_packed=5 if pack else 0)


@pytest.mark.parametrize("lds", [16 * 1024, 24 * 1024, 152 * 1024])
def test_sim_count_kernel_lds_sizes(lds):
    # 40 targets over ranges of 330-355 values: read as rows and values, 3
    # target slices per column in 16 KiB, 2 in 24 KiB, one workgroup per column
    # in 152 KiB; as packed entries the bins are only the values that occur
    # (fewer), so at most as many workgroups
    d, r, ip, g = _rw(22, 6000, 4, 40, 0.6, lambda rg, n: rg.negative_binomial(0.8, 0.02, n) + 1)
    blocks = {16 * 1024: 12, 24 * 1024: 8, 152 * 1024: 4}[lds]
    run_count(d, r, ip, g, 40, 6000, "min", lds, expect_counted=4, expect_blocks=blocks,
              pack=False)
    run_count(d, r, ip, g, 40, 6000, "min", lds, expect_counted=4, expect_packed=None)
    assert 4 <= sim.mwu.last_blocks <= blocks and sim.mwu.last_packed == sim.mwu.last_blocks


def test_sim_count_kernel_hands_over_wide_ranges():
    # values span more than hist_max bins: read as rows and values every column
    # goes to the split path; as packed entries the bins are the ~1,200
    # distinct values a column holds, and the count kernel takes them
    d, r, ip, g = _rw(23, 3000, 3, 5, 0.4, lambda rg, n: rg.integers(-20000, 20000, n))
    run_count(d, r, ip, g, 5, 3000, "mix", 76 * 1024, expect_counted=0, pack=False)
    run_count(d, r, ip, g, 5, 3000, "mix", 76 * 1024, expect_counted=3, expect_packed=3)
    # a low histogram limit hands over a narrow range too
    d, r, ip, g = _rw(24, 3000, 3, 5, 0.4, lambda rg, n: rg.integers(1, 300, n))
    run_count(d, r, ip, g, 5, 3000, "min", 76 * 1024, hist_max=128, expect_counted=0)


def test_sim_count_kernel_too_many_slices_take_split_path():
    # 300 targets x 1,000 values in 16 KiB: a workgroup holds 3 targets' rows,
    # 100 slices per column (more than 64): the columns take the split path
    d, r, ip, g = _rw(25, 4000, 2, 300, 0.5, lambda rg, n: rg.integers(1, 1001, n))
    run_count(d, r, ip, g, 300, 4000, "min", 16 * 1024, expect_counted=0, expect_blocks=0)


def test_sim_count_kernel_nonfinite_and_mixed_columns():
    d, r, ip, g = _rw(26, 2500, 4, 4, 0.4, lambda rg, n: rg.integers(0, 6, n))
    d = d.copy()
    c0 = slice(ip[0], ip[1])
    d[c0][::7] = np.nan          # column 0: NaNs are skipped (zero block outside none)
    d[ip[1]:ip[2]][::11] = np.inf
    d[ip[2]] = 0.5               # column 2: one non-integer -> split path
    d[ip[3] + 2] = -np.inf
    for mode in ("min", "none"):
        for pack in (True, False):   # non-finite values take the skip code when packed
            run_count(d, r, ip, g, 4, 2500, mode, 76 * 1024, expect_counted=3, pack=pack,
                      expect_packed=3 if pack else 0)


def test_sim_count_kernel_packed_rows_out_of_range_and_sizes():
    # R = 300,000 leaves 13 code bits beside a row: a column of 8,190 bins is
    # packed, one of 8,191 (count path, read as rows and values) is not; rows
    # outside [0, R) are packed as R (no group)
    rng = np.random.default_rng(37)
    R, T = 300_000, 3
    rows, ip = synthetic._bernoulli_rows(rng, R, np.array([0.03, 0.03, 0.01]))
    rows = rows.copy()
    d = rng.integers(1, 8190, len(rows)).astype(np.float32)
    d[ip[0]], d[ip[0] + 1] = 1.0, 8190.0        # column 0: 8,190 bins
    d[ip[1]], d[ip[1] + 1] = 1.0, 8191.0        # column 1: 8,191 bins
    d[ip[2]:ip[3]] = rng.integers(0, 9, ip[3] - ip[2])
    rows[ip[2] + 5] = -4
    rows[ip[2] + 9] = R + 7
    gid = rng.integers(-1, T + 1, R).astype(np.int32)
    for mode in ("min", "none"):
        run_count(d, rows, ip, gid, T, R, mode, 152 * 1024, expect_counted=3, expect_blocks=3,
                  expect_packed=2)


def test_sim_count_kernel_rows_beyond_24_bits():
    # R = 2^24 + 100: label positions take the 32-bit multiply; 25 row bits
    # leave 7 code bits, so the 100-bin column is packed and the 300-bin one
    # is read as rows and values; rows near R
    rng = np.random.default_rng(38)
    R, T = (1 << 24) + 100, 4
    n = 6000
    rows0 = np.sort(rng.choice(R, n, replace=False))
    rows0[-3:] = [R - 3, R - 2, R - 1]
    rows1 = np.sort(rng.choice(R, n, replace=False))
    rows = np.concatenate([rows0, rows1]).astype(np.int64)
    ip = np.array([0, n, 2 * n], dtype=np.int64)
    d = np.concatenate([rng.integers(1, 101, n), rng.integers(1, 301, n)]).astype(np.float32)
    d[0], d[1] = 1.0, 100.0
    d[n], d[n + 1] = 1.0, 300.0
    gid = np.full(R, -1, np.int32)
    pick = rng.choice(R, 400_000, replace=False)
    gid[pick] = rng.integers(0, T + 1, len(pick))
    gid[rows] = rng.integers(0, T + 1, len(rows))
    for mode in ("min", "none"):
        run_count(d, rows, ip, gid, T, R, mode, 76 * 1024, expect_counted=2, expect_blocks=2,
                  expect_packed=1)


def test_sim_count_kernel_extreme_values():
    # values at the ends of the 16-bit range the column metadata admits:
    # -32768 .. -32700 in column 0, 32700 .. 32767 in column 1, both ends
    # (a 65,536-value range: split path) in column 2
    d, r, ip, g = _rw(30, 3000, 3, 6, 0.4, lambda rg, n: rg.integers(0, 69, n))
    d = d.copy()
    d[ip[0]:ip[1]] = np.where(d[ip[0]:ip[1]] == 0, 0, d[ip[0]:ip[1]] - 32768)
    d[ip[0]] = -32768
    d[ip[1]:ip[2]] = np.where(d[ip[1]:ip[2]] == 0, 0, d[ip[1]:ip[2]] + 32699)
    d[ip[1]] = 32767
    d[ip[2]] = -32768
    d[ip[2] + 1] = 32767
    for mode in ("min", "mix", "none"):
        # rows and values: column 2's 65,536-value range takes the split path;
        # packed, its ~70 distinct values are the bins and it is counted too
        run_count(d, r, ip, g, 6, 3000, mode, 76 * 1024, expect_counted=2, expect_blocks=2,
                  pack=False)
        run_count(d, r, ip, g, 6, 3000, mode, 76 * 1024, expect_counted=3, expect_blocks=3,
                  expect_packed=3)


@pytest.mark.parametrize("mixed", [False, True])
def test_sim_count_kernel_metadata_mismatch_is_an_error(mixed):
    # recorded ranges one value short of the data: every counted column meets
    # a value outside its range, and the job reports the status (naming the
    # lowest such column) instead of counting it wrongly; a column the host
    # routed to the sort path beside them is still done.  The columns are
    # read as rows and values: packed entries rank the values a column holds,
    # so they cannot meet one outside their bins.
    d, r, ip, g = _rw(32, 3000, 5, 6, 0.5, lambda rg, n: rg.integers(-3, 12, n))
    if mixed:
        d = d.copy()
        d[ip[4]] = 0.5          # column 4 is not integer-valued: routed by the host
    U1, P, TIE, n1, n2, err, status = sim.mwu(d, r, ip, g, 6, 3000, mode=ZH["min"],
                                              shrink_range=True, timeout_s=120, pack=False)
    assert status == 4 and err == 2 ** 64 - 1   # kStatusRange
    assert sim.mwu.last_counted == (4 if mixed else 5)
    if mixed:
        oU1, oP, oT, on1, on2 = oracle.mwu(d, r, ip, g, 6, n_rows=3000, zero_handling="min",
                                           use_continuity=True, with_stats=True)
        assert np.array_equal(U1[24:30], oU1[24:30]) and np.array_equal(TIE[24:30], oT[24:30])


def test_sim_count_kernel_eight_byte_values():
    d, r, ip, g = _rw(27, 2000, 3, 4, 0.4, lambda rg, n: rg.integers(-30, 200, n))
    run_count(d.astype(np.float64), r, ip, g, 4, 2000, "mix", 76 * 1024, f64=True,
              expect_counted=3)


@pytest.mark.parametrize("lds", [76 * 1024, 24 * 1024])
def test_sim_count_kernel_target_slices(lds):
    # 200 targets x 400 values do not fit one workgroup: 76 KiB holds 88
    # targets' rows (3 slices of 67), 24 KiB 24 (9 slices of 23)
    d, r, ip, g = _rw(28, 8000, 3, 200, 0.5, lambda rg, n: rg.integers(1, 401, n))
    blocks = 3 * (3 if lds == 76 * 1024 else 9)
    for mode, block in (("min", 512), ("mix", 256), ("none", 512)):
        run_count(d, r, ip, g, 200, 8000, mode, lds, expect_counted=3, block=block,
                  expect_blocks=blocks)


def test_sim_count_kernel_wide_counters():
    # a column of more than 65,535 entries: 32-bit counters, and a 3,000-value
    # range of which a 76 KiB workgroup holds 3 targets' rows (7 slices)
    d, r, ip, g = _rw(29, 70000, 1, 20, 0.96, lambda rg, n: rg.integers(1, 3001, n))
    assert ip[1] >= 65536
    run_count(d, r, ip, g, 20, 70000, "min", 76 * 1024, expect_counted=1, expect_blocks=7)


@pytest.mark.parametrize("mode,f64", [("min", False), ("none", True)], ids=["min-f32", "none-f64"])
def test_sim_reference_beyond_chip(mode, f64):
    # reference slices above the on-chip sort (8,192 four-byte keys, 4,096
    # eight-byte keys): ~10,000 values per column sorted in HBM (two or three
    # LDS-sorted runs and merge passes through the spill space), ranked from
    # HBM; ties inside the reference (values at 0.01) and against targets,
    # negatives, and a column whose reference slice fits the chip beside them
    rng = np.random.default_rng(35)
    R, T = 14000, 4
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.95, 0.8, 0.3]))
    data = np.round(rng.normal(0, 3, len(rows)), 2)
    data = data.astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:11000]] = 0
    gid[perm[11000:]] = rng.integers(1, T + 1, R - 11000)
    run_both(data, rows, indptr, gid, T, R, mode, f64=f64)


@pytest.mark.parametrize("cap,f64", [(64, False), (256, False), (128, True)],
                         ids=["cap64-f32", "cap256-f32", "cap128-f64"])
def test_sim_reference_cutover(cap, f64):
    # the LDS/global sort cutover as a tuning (ref_lds_keys): reference
    # slices of a few hundred values, above a low cutover, sorted in HBM and
    # ranked through it, beside slices at or below it ranked from LDS; ties,
    # negatives and targets both smaller and larger than 64 keys
    rng = np.random.default_rng(41)
    R, T = 3000, 12
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.9, 0.5, 0.2, 0.05]))
    data = np.round(rng.normal(0, 2, len(rows)), 1).astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:700]] = 0
    gid[perm[700:]] = rng.integers(1, T + 1, R - 700)
    for mode in ("min", "none"):
        run_both(data, rows, indptr, gid, T, R, mode, f64=f64, ref_cap=cap)


@pytest.mark.parametrize("cap,f64,modes", [(8192, False, ("min",)), (128, False, ("mix", "none")),
                                           (128, True, ("min",))],
                         ids=["onchip-f32", "cap128-f32", "cap128-f64"])
def test_sim_kept_reference_tables(cap, f64, modes):
    # a job run, its reference tables kept by the fill kernel, then the job
    # run again with its own tables wiped: every sort-path column ranks from
    # the kept sorted reference (on chip and beyond the cutover), bit-identical
    # to the oracle; integer columns on the histogram branch are not kept
    rng = np.random.default_rng(47)
    R, T = 1500, 10
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.9, 0.5, 0.2, 0.05, 0.4]))
    data = np.round(rng.normal(0, 2, len(rows)), 1)
    last = slice(indptr[4], indptr[5])
    data[last] = rng.integers(1, 30, indptr[5] - indptr[4])   # an integer column
    data = data.astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:500]] = 0
    gid[perm[500:]] = rng.integers(1, T + 1, R - 500)
    for mode in modes:
        run_both(data, rows, indptr, gid, T, R, mode, f64=f64, ref_cap=cap, count_lds=0,
                 refcache=True)
        assert sim.mwu.last_cached == 4


@pytest.mark.parametrize("f64,modes", [(False, ("min", "none")), (True, ("mix",))],
                         ids=["f32", "f64"])
def test_sim_split_with_counters_in_memory(f64, modes):
    # the split of a job with more groups than LDS holds counters for: counts
    # by device-scope atomics in the job's arrays, offsets scanned there, keys
    # to their slots (a dense column) or offsets (sparse ones), an integer
    # column on the histogram branch; bit-identical to the oracle
    rng = np.random.default_rng(53)
    R, T = 1200, 20
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([1.0, 0.6, 0.35, 0.5]))
    data = np.round(rng.normal(0, 2, len(rows)), 1)
    data[indptr[3]:indptr[4]] = rng.integers(-3, 30, indptr[4] - indptr[3])
    data = data.astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:400]] = 0
    gid[perm[400:1100]] = rng.integers(1, T + 1, 700)
    for mode in modes:
        run_both(data, rows, indptr, gid, T, R, mode, f64=f64, count_lds=0, global_split=True)


@pytest.mark.parametrize("f64,mode", [(False, "min"), (False, "none"), (True, "mix")],
                         ids=["f32-min", "f32-none", "f64-mix"])
def test_sim_dense_columns_through_the_member_list(f64, mode):
    # columns storing rows 0 .. R-1 in order gather their groups' rows through
    # the job's member list instead of reading every entry: with explicit
    # zeros, NaN and an integer column (histogram branch), slots staged in LDS
    # and -- for the wide groups -- written to HBM, the results equal the
    # oracle's and those of the entry-by-entry read
    rng = np.random.default_rng(61)
    R, C, T = 3000, 3, 60
    rows, indptr = _dense(R, C)
    data = np.round(rng.normal(0, 1.5, R * C), 1)
    z = rng.random(R * C)
    data[z < 0.08] = 0.0
    data[(z >= 0.08) & (z < 0.1)] = np.nan
    data[2 * R:] = rng.integers(0, 12, R)   # an integer column
    data = data.astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:400]] = 0
    at = 400
    for g, n in enumerate(rng.integers(3, 30, T), start=1):
        gid[perm[at:at + n]] = g
        at += n
    run_both(data, rows, indptr, gid, T, R, mode, f64=f64, count_lds=0)
    a = sim.mwu(data, rows, indptr, gid, T, R, mode=ZH[mode], f64=f64, count_lds=0,
                timeout_s=120)
    b = sim.mwu(data, rows, indptr, gid, T, R, mode=ZH[mode], f64=f64, count_lds=0,
                timeout_s=120, members=False)
    for x, y in zip(a, b):
        assert np.array_equal(x, y, equal_nan=True)


def _fullest_bucket(keys):
    """Keys in the fullest bucket of an on-chip reference's index: buckets
    the power of two >= its size (64 to 4,096), the key range shifted onto
    them -- computed here only to show which sort a column takes."""
    na = len(keys)
    if na == 0:
        return 0
    nb = 64
    while nb < na and nb < 4096:
        nb *= 2
    k = keys.astype(np.uint64)
    lo, hi = int(k.min()), int(k.max())
    shift = max((hi - lo).bit_length() - (nb.bit_length() - 1), 0)
    return int(np.bincount(((k - np.uint64(lo)) >> np.uint64(shift)).astype(np.int64)).max())


def _sort_keys(v):
    """Order-preserving unsigned keys of floats (negatives flipped)."""
    b = v.view(np.uint64 if v.dtype == np.float64 else np.uint32)
    top = np.array(1, b.dtype) << np.array(8 * b.itemsize - 1, b.dtype)
    return np.where(b & top, ~b, b | top)


@pytest.mark.parametrize("f64", [False, True], ids=["f32", "f64"])
def test_sim_reference_sort_by_buckets_and_bitonic(f64):
    # on-chip references sorted by buckets (the fullest bucket holds at most
    # 64 keys: distinct values, and values at 0.01 with runs of equal keys)
    # and by the bitonic sort (a handful of distinct values: full buckets),
    # in every zero mode; the bucket sort's tie terms come from the runs
    rng = np.random.default_rng(47)
    R, T = 5000, 6
    dens = np.array([0.6, 0.6, 0.6, 0.2])
    rows, indptr = synthetic._bernoulli_rows(rng, R, dens)
    n = np.diff(indptr)
    # (8-byte keys of values on both sides of zero span some 2,000 binades:
    # one bucket each, too coarse for the bucket sort, so positive values there)
    sign = (lambda x: np.abs(x) + 0.01) if f64 else (lambda x: x)
    parts = [sign(rng.normal(0, 2, n[0])), sign(np.round(rng.normal(0, 2, n[1]), 2)),
             rng.choice([-1.5, 0.25, 0.75, 2.5], n[2]), np.log1p(rng.lognormal(0, 1.5, n[3]))]
    data = np.concatenate(parts).astype(np.float64 if f64 else np.float32)
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:1800]] = 0
    gid[perm[1800:]] = rng.integers(1, T + 1, R - 1800)
    fullest = []
    for c in range(len(n)):
        r, v = rows[indptr[c]:indptr[c + 1]], data[indptr[c]:indptr[c + 1]]
        fullest.append(_fullest_bucket(_sort_keys(v[(gid[r] == 0) & (v != 0)])))
    assert max(fullest[0], fullest[1], fullest[3]) <= 64 < fullest[2]
    for mode in ("min", "max", "mix", "none"):
        run_both(data, rows, indptr, gid, T, R, mode, f64=f64)


@pytest.mark.parametrize("mode", ["min", "max", "mix", "none"])
def test_sim_sorted_pass_hashed_and_sorted_targets(mode):
    # targets of 65 to 192 keys (ranked in chunks of 64 with a tie table, no
    # sort) beside targets of more (sorted) and of at most 64 (windows):
    # many equal keys inside a target (values at 0.1), negatives, zeros
    rng = np.random.default_rng(53)
    R = 4000
    rows, indptr = synthetic._bernoulli_rows(rng, R, np.array([0.9, 0.6, 0.35]))
    data = np.round(rng.normal(0, 1.5, len(rows)), 1).astype(np.float32)
    sizes = [90, 120, 140, 180, 260, 30, 10, 400]
    gid = np.full(R, -1, np.int32)
    perm = rng.permutation(R)
    gid[perm[:1200]] = 0
    k = 1200
    for t, n in enumerate(sizes, start=1):
        gid[perm[k:k + n]] = t
        k += n
    keys = [[int(((gid[rows[indptr[c]:indptr[c + 1]]] == t) &
                  (data[indptr[c]:indptr[c + 1]] != 0)).sum()) for t in range(1, len(sizes) + 1)]
            for c in range(3)]
    flat = [x for col in keys for x in col]
    assert any(64 < x <= 128 for x in flat) and any(128 < x <= 192 for x in flat)
    assert any(x > 192 for x in flat)
    run_both(data, rows, indptr, gid, len(sizes), R, mode)

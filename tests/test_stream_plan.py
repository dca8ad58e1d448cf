"""Gene-run planning of the streaming driver (CPU): runs cover every gene once,
in order, each within its budget; a gene larger than the budget raises
MemoryError, as the reference's planner does for a group (stream.py:188-206)."""
import numpy as np
import pytest

from hpdex_amd import stream


def test_runs_cover_genes_within_budget():
    rng = np.random.default_rng(1)
    nnz = rng.integers(0, 5000, 700)
    indptr = np.concatenate([[0], np.cumsum(nnz)])
    T = 40
    budget = stream.run_bytes(60_000, 30, T)
    runs = stream.plan_gene_runs(indptr, T, budget)
    assert runs[0][0] == 0 and runs[-1][1] == 700
    assert all(a[1] == b[0] for a, b in zip(runs, runs[1:]))
    for c0, c1 in runs:
        assert c1 > c0
        assert stream.run_bytes(int(indptr[c1] - indptr[c0]), c1 - c0, T) <= budget
    # greedy: no run could have taken the next gene as well
    for (c0, c1), (_, _) in zip(runs, runs[1:]):
        assert stream.run_bytes(int(indptr[c1 + 1] - indptr[c0]), c1 + 1 - c0, T) > budget


def test_gene_larger_than_budget_raises():
    indptr = np.array([0, 10, 10_000_000, 10_000_010])
    with pytest.raises(MemoryError, match="Gene 1"):
        stream.plan_gene_runs(indptr, 5, stream.run_bytes(1_000_000, 1, 5))


def test_empty_matrix_has_no_runs():
    assert stream.plan_gene_runs(np.array([0]), 3, 1 << 20) == []


def test_run_bytes_counts_the_library_footprint_of_near_empty_genes():
    """A run's cost includes what the library allocates per gene whatever its
    size (the sort path's reference table, column plan, count-kernel items,
    per-group counters) and per run (the labels of every row): for genes of
    one stored value each, 1,000 genes over 150,000 cells take more than
    1,000 reference tables of 32 KiB plus the 600 KB of labels."""
    from hpdex_amd import _capi
    R, T = 150_000, 4
    dev = _capi.de_csc_device_bytes(np.float32, R, 1000, 1000, T)
    assert dev >= 1000 * 32_772 + R * 4
    assert stream.run_bytes(1000, 1000, T, R) >= dev
    # the run grows with its genes, values and targets
    assert _capi.de_csc_device_bytes(np.float32, R, 1001, 1000, T) > dev
    assert _capi.de_csc_device_bytes(np.float32, R, 1000, 5000, T) > dev
    assert _capi.de_csc_device_bytes(np.float32, R, 1000, 1000, T + 1) > dev
    assert _capi.de_csc_device_bytes(np.float64, R, 1000, 1000, T) > dev
    # near-empty genes: runs are cut by the per-gene cost, not by their values
    indptr = np.arange(2001, dtype=np.int64)
    budget = stream.run_bytes(500, 500, T, R)
    runs = stream.plan_gene_runs(indptr, T, budget, R)
    assert max(c1 - c0 for c0, c1 in runs) == 500

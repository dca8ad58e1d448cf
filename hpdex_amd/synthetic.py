"""Seeded synthetic single-cell workloads (cells x genes CSC, float32).

Each generator returns a dict with ``data, indices, indptr`` (CSC, int64
indices), ``n_rows``, ``n_cols``, ``group_id`` (int32, 0 = reference,
1..T targets, -1 ignored), ``n_targets`` and the options the workload runs
with.  Seeds are fixed per workload (1000 + workload number).

Rules the workload descriptions leave open, stated here:
  * nonzero positions are a Bernoulli(density) process per gene (independent
    rows), drawn as geometric gaps;
  * target sizes that would need more cells than exist are rescaled
    proportionally (floor, then clipped back to the lower bound), and if the
    total still exceeds the cells available the largest targets are trimmed.
"""
from __future__ import annotations

import os

import numpy as np


def _bernoulli_rows(rng, n_rows, dens):
    """Sorted distinct row indices per column, each row kept with prob dens[c]."""
    dens = np.clip(np.asarray(dens, dtype=np.float64), 1e-6, 1.0)
    C = len(dens)
    budget = np.ceil(n_rows * dens * 1.25 + 64).astype(np.int64)
    gaps = rng.geometric(np.repeat(dens, budget)).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(budget)])
    cs = np.cumsum(gaps)
    seg_base = np.repeat(np.concatenate([[0], cs[starts[1:-1] - 1]]), budget)
    pos = cs - seg_base - 1
    col = np.repeat(np.arange(C, dtype=np.int64), budget)
    keep = pos < n_rows
    # columns whose gap budget ran out before the last row: extend them
    last = pos[starts[1:] - 1]
    short = np.nonzero(last < n_rows)[0]
    rows_list = [pos[keep]]
    cols_list = [col[keep]]
    for c in short:
        p = int(last[c])
        extra = []
        while True:
            p += int(rng.geometric(dens[c]))
            if p >= n_rows:
                break
            extra.append(p)
        if extra:
            rows_list.append(np.asarray(extra, dtype=np.int64))
            cols_list.append(np.full(len(extra), c, dtype=np.int64))
    rows = np.concatenate(rows_list)
    cols = np.concatenate(cols_list)
    if len(short):
        o = np.lexsort((rows, cols))
        rows, cols = rows[o], cols[o]
    indptr = np.zeros(C + 1, dtype=np.int64)
    np.cumsum(np.bincount(cols, minlength=C), out=indptr[1:])
    return rows, indptr


def _nb(rng, mean, r, size):
    mean = np.asarray(mean, dtype=np.float64)
    return rng.negative_binomial(r, r / (r + mean), size=size)


def _target_sizes(rng, T, median, sigma, lo, hi, avail):
    s = np.clip(np.floor(rng.lognormal(np.log(median), sigma, T)), lo, hi).astype(np.int64)
    if s.sum() > avail:
        s = np.maximum(lo, np.floor(s * (avail / s.sum())).astype(np.int64))
        while s.sum() > avail:
            i = int(np.argmax(s))
            s[i] -= min(s[i] - lo, s.sum() - avail) if s[i] > lo else 0
            if s[i] <= lo and s.sum() > avail:
                raise ValueError("target sizes cannot fit the available cells")
    return s


def _assign_groups(rng, n_rows, n_ref, sizes):
    gid = np.full(n_rows, -1, dtype=np.int32)
    perm = rng.permutation(n_rows)
    gid[perm[:n_ref]] = 0
    off = n_ref
    for t, sz in enumerate(sizes, start=1):
        gid[perm[off:off + sz]] = t
        off += sz
    return gid


def small_case(seed: int = 1001):
    """2,000 cells x 500 genes, 10% nonzeros at uniform positions,
    values NB(mean 3, r=0.5)+1; 1,000 reference cells and 4 targets of 250."""
    rng = np.random.default_rng(seed)
    R, C = 2000, 500
    flat = rng.choice(R * C, size=R * C // 10, replace=False)
    flat.sort()
    cols, rows = np.divmod(flat, R)   # column-major positions
    vals = (_nb(rng, 3.0, 0.5, len(flat)) + 1).astype(np.float32)
    indptr = np.zeros(C + 1, dtype=np.int64)
    np.cumsum(np.bincount(cols, minlength=C), out=indptr[1:])
    gid = np.concatenate([np.zeros(1000), np.repeat(np.arange(1, 5), 250)]).astype(np.int32)
    rng.shuffle(gid)
    return dict(name="small", data=vals, indices=rows.astype(np.int64), indptr=indptr,
                n_rows=R, n_cols=C, group_id=gid, n_targets=4,
                opts=dict(tie_correction=True, use_continuity=True, zero_handling="min",
                          method="asymptotic"))


def histogram_case(seed: int = 1002, n_cells: int = 150_000, n_genes: int = 8_000,
                   n_ref: int = 15_000, n_targets: int = 300):
    """Integer UMI counts as float32: per-gene density ~ Beta(1,12), per-gene
    NB mean ~ LogNormal(0,1), r=0.8, values NB+1 capped at 4,095; reference
    15,000 cells; 300 targets of LogNormal(median 300, sigma 0.8) cells clipped
    to [20, 3,000]; remaining cells ignored (-1)."""
    rng = np.random.default_rng(seed)
    dens = rng.beta(1.0, 12.0, n_genes)
    mu = rng.lognormal(0.0, 1.0, n_genes)
    rows, indptr = _bernoulli_rows(rng, n_cells, dens)
    nnz_per = np.diff(indptr)
    vals = _nb(rng, np.repeat(mu, nnz_per), 0.8, len(rows)) + 1
    np.minimum(vals, 4095, out=vals)
    sizes = _target_sizes(rng, n_targets, 300, 0.8, 20, 3000, n_cells - n_ref)
    gid = _assign_groups(rng, n_cells, n_ref, sizes)
    return dict(name="histogram", data=vals.astype(np.float32), indices=rows, indptr=indptr,
                n_rows=n_cells, n_cols=n_genes, group_id=gid, n_targets=n_targets,
                opts=dict(tie_correction=True, use_continuity=True, zero_handling="min",
                          method="asymptotic"))


def float_case(seed: int = 1003, n_cells: int = 400_000, n_genes: int = 18_000,
               n_ref: int = 40_000, n_targets: int = 1000, density: float = 0.05,
               zero_handling: str = "min"):
    """log1p(CP10k) floats from NB(LogNormal(0,1.2) mean, r=0.6) raw counts at
    5% density (raw zeros inside the pattern are kept as explicit zeros);
    reference 40,000 cells; 1,000 targets of LogNormal(median 250, sigma 1.0)
    clipped to [10, 5,000]."""
    rng = np.random.default_rng(seed)
    mu = rng.lognormal(0.0, 1.2, n_genes)
    rows, indptr, raw = _counts_matrix_parallel(int(rng.integers(1 << 62)), n_cells,
                                                np.full(n_genes, density), mu, 0.6,
                                                plus_one=False)
    lib = np.bincount(rows, weights=raw, minlength=n_cells)
    lib[lib == 0] = 1.0
    vals = np.log1p(raw / lib[rows] * 1e4).astype(np.float32)
    sizes = _target_sizes(rng, n_targets, 250, 1.0, 10, 5000, n_cells - n_ref)
    gid = _assign_groups(rng, n_cells, n_ref, sizes)
    return dict(name="float", data=vals, indices=rows, indptr=indptr, n_rows=n_cells,
                n_cols=n_genes, group_id=gid, n_targets=n_targets,
                opts=dict(tie_correction=True, use_continuity=True, zero_handling=zero_handling,
                          method="asymptotic"))


def exact_case(seed: int = 1004, n_cells: int = 150_000, n_genes: int = 6_000,
               n_ref: int = 60, n_targets: int = 2000):
    """Dense N(0,1) float32 plus a per-(target, gene) N(0,0.5) shift on 20% of
    genes; reference 60 cells; 2,000 targets of U[3,25] cells; exact p."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(3, 26, n_targets)
    gid = _assign_groups(rng, n_cells, n_ref, sizes)
    X = rng.standard_normal((n_genes, n_cells), dtype=np.float32)   # gene-major = CSC data
    shifted = np.nonzero(rng.random(n_genes) < 0.2)[0]
    shift = rng.normal(0.0, 0.5, (len(shifted), n_targets + 1)).astype(np.float32)
    shift[:, 0] = 0.0
    sel = gid >= 0
    cells = np.nonzero(sel)[0]
    for k, g in enumerate(shifted):
        X[g, cells] += shift[k, gid[cells]]
    data = X.reshape(-1)
    indices = np.tile(np.arange(n_cells, dtype=np.int64), n_genes)
    indptr = np.arange(n_genes + 1, dtype=np.int64) * n_cells
    return dict(name="exact", data=data, indices=indices, indptr=indptr, n_rows=n_cells,
                n_cols=n_genes, group_id=gid, n_targets=n_targets,
                opts=dict(tie_correction=True, use_continuity=False, zero_handling="min",
                          method="exact"))


def subset_columns(w: dict, cols) -> dict:
    """The same workload restricted to the given gene columns (for bounded
    CPU checks)."""
    cols = np.asarray(cols, dtype=np.int64)
    ip = w["indptr"]
    lens = ip[cols + 1] - ip[cols]
    idx = np.concatenate([np.arange(ip[c], ip[c + 1]) for c in cols]) if len(cols) else \
        np.zeros(0, dtype=np.int64)
    new_ip = np.zeros(len(cols) + 1, dtype=np.int64)
    np.cumsum(lens, out=new_ip[1:])
    out = dict(w)
    out.update(data=w["data"][idx], indices=w["indices"][idx], indptr=new_ip, n_cols=len(cols))
    return out


def _counts_matrix_parallel(seed, R, dens, mu, r, n_chunks=64, threads=None, plus_one=True):
    """CSC of NB(mu[c], r) (+ 1) counts at Bernoulli(dens[c]) positions, built
    over n_chunks fixed column chunks, each with its own generator spawned from
    `seed` (numpy's samplers release the GIL, so the chunks run on threads; the
    result does not depend on the thread count)."""
    from concurrent.futures import ThreadPoolExecutor
    C = len(dens)
    bounds = np.linspace(0, C, min(n_chunks, max(C, 1)) + 1).astype(np.int64)
    seqs = np.random.SeedSequence(seed).spawn(len(bounds) - 1)

    def chunk(k):
        g = np.random.default_rng(seqs[k])
        a, b = bounds[k], bounds[k + 1]
        rows, ip = _bernoulli_rows(g, R, dens[a:b])
        vals = _nb(g, np.repeat(mu[a:b], np.diff(ip)), r, len(rows))
        vals = (vals + 1).astype(np.float32) if plus_one else vals.astype(np.float32)
        return rows, ip, vals

    n = threads or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(n) as ex:
        parts = list(ex.map(chunk, range(len(bounds) - 1)))
    lens = [len(p[0]) for p in parts]
    offs = np.concatenate([[0], np.cumsum(lens)])
    indptr = np.concatenate([[0]] + [p[1][1:] + offs[k] for k, p in enumerate(parts)])
    rows = np.concatenate([p[0] for p in parts])
    vals = np.concatenate([p[2] for p in parts])
    return rows, indptr.astype(np.int64), vals


def serving_matrix(rng):
    """One matrix of the serving distribution: cells U[5k, 200k], genes
    U[2k, 20k], density U[3%, 12%] (one density per matrix), per-gene NB mean
    ~ LogNormal(0, 1), r = 0.7, values NB + 1 as float32."""
    R = int(rng.integers(5_000, 200_001))
    C = int(rng.integers(2_000, 20_001))
    dens = float(rng.uniform(0.03, 0.12))
    mu = rng.lognormal(0.0, 1.0, C)
    rows, indptr, vals = _counts_matrix_parallel(int(rng.integers(1 << 62)), R, np.full(C, dens),
                                                 mu, 0.7)
    return dict(data=vals, indices=rows, indptr=indptr, n_rows=R, n_cols=C, density=dens)


def serving_workload(seed: int = 1005, n_matrices: int = 4, n_jobs: int = 64,
                     resident_frac: float = 0.3):
    """The serving workload: a pool of matrices drawn from the serving
    distribution and a job list over it.  Each job takes a pool matrix, a
    reference of 10% of its cells and 5-500 targets of LogNormal(median 200,
    sigma 0.9) cells (sizes rescaled to fit the remaining cells); a fraction
    `resident_frac` of the jobs re-query the matrix resident in HBM with a new
    target set, the others bring it as a new host matrix."""
    rng = np.random.default_rng(seed)
    pool = [serving_matrix(rng) for _ in range(n_matrices)]
    jobs = []
    for _ in range(n_jobs):
        k = int(rng.integers(0, n_matrices))
        R = pool[k]["n_rows"]
        n_ref = R // 10
        T = int(rng.integers(5, 501))
        # at least two cells per target: a group of one has no Mann-Whitney
        # test ("Sample too small", src/cpp/mannwhitneyu.cpp:980-984)
        sizes = _target_sizes(rng, T, 200, 0.9, 2, R, R - n_ref)
        gid = _assign_groups(rng, R, n_ref, sizes)
        jobs.append(dict(matrix=k, group_id=gid, n_targets=T,
                         resident=bool(rng.random() < resident_frac)))
    return pool, jobs


def algorithmic_bytes(w: dict, value_bytes: int = 4) -> int:
    """nnz*(value+4-byte row) + (C+1)*8 + R*4 + C*T*16: the bytes a single
    pass over the job must move (inputs once, U1 and P out)."""
    nnz = int(w["indptr"][-1])
    C, R, T = w["n_cols"], w["n_rows"], w["n_targets"]
    return pass_bytes(nnz, C, R, T, value_bytes)


def pass_bytes(nnz: int, C: int, R: int, T: int, value_bytes: int = 4) -> int:
    """Algorithmic bytes of one pass over C columns holding nnz stored values:
    values and row indices once, the column pointers, the group label of every
    row, and U1 and P (8 bytes each) of every (column, target) pair."""
    return int(nnz) * (value_bytes + 4) + (int(C) + 1) * 8 + int(R) * 4 + int(C) * int(T) * 16

"""Autotuner for the device path's tunable choices.

The reference has one fixed CPU schedule (OpenMP over genes,
src/cpp/mannwhitneyu.cpp:1180-1304); on the device the same job can be laid
out several ways, and which is fastest depends on the matrix:

* ``hist_max_bins``  -- the cutover between the one-pass histogram kernel and
  the split / reference-sort / target kernels: integer columns of at most this
  many bins take the histogram kernel, every other column is sorted
  (0 sends every column to the sort path);
* ``count_lds_bytes`` -- LDS per histogram workgroup: more LDS holds more
  target rows per workgroup (fewer slices re-reading a column) at fewer
  workgroups per CU;
* ``count_block``    -- threads per histogram workgroup (256, 512, 1024);
* ``split_arrival``  -- sort path: the split kernel keeps a sparse column's
  entry groups in LDS and reads only its values again to scatter them (1), or
  reads rows, values and labels twice (0);
* ``ref_lds_keys``   -- sort path, the LDS/global sort cutover: a reference
  slice of at most this many values is sorted in LDS and ranked from LDS, a
  larger one is sorted in HBM (merge passes) and ranked through L2.

Candidates that lay the job out the same way as one already timed (the
histogram knobs of a job with no histogram column, the sort-path knobs of a job
with no sort-path column) are not timed again: they take that trial's time.

``autotune`` times every candidate on the caller's own resident matrix and
group labels in the batched (graph-replayed) form, checks that every
candidate's U1 and p equal the first candidate's bit for bit (the layout
changes the schedule, never the arithmetic), applies the fastest and returns
the trial table.  ``save`` / ``load`` keep the choice in a JSON file so a
serving process can apply it at start-up without timing again.
"""
from __future__ import annotations

import itertools
import json

import numpy as np

from . import _capi

# Searched by default: the histogram/sort cutover, LDS per histogram workgroup
# (four, three or one workgroup per CU; 40,448 B is the built-in default) and
# the histogram workgroup size.
DEFAULT_SPACE = {
    "hist_max_bins": (8192, 2048, 0),
    "count_lds_bytes": (40448, 52 * 1024, 152 * 1024),
    "count_block": (256, 512, 1024),
    "split_arrival": (1, 0),
    "ref_lds_keys": (8192, 2048, 512),
}
COUNT_KNOBS = ("count_lds_bytes", "count_block", "count_packed")
SORT_KNOBS = ("split_arrival", "ref_lds_keys")


def space_for(matrix=None) -> dict:
    """The search space for a job on `matrix` (every knob of DEFAULT_SPACE)."""
    return dict(DEFAULT_SPACE)

KNOBS = ("hist_max_bins", "size_ordered", "count_lds_bytes", "count_block", "count_packed",
         "split_arrival", "ref_lds_keys", "ref_cache")


class TuningMismatch(_capi.HpdexError):
    """Two candidate tunings gave different results for the same job."""


def current() -> dict:
    """The tuning in force, as a dict of the ``hpdex_tuning`` fields."""
    t = _capi.get_tuning()
    return {k: int(getattr(t, k)) for k in KNOBS}


def apply(tuning: dict) -> None:
    """Set the fields named in ``tuning`` (others keep their value)."""
    _capi.set_tuning(**{k: int(v) for k, v in tuning.items() if k in KNOBS})


def candidates(space: dict | None = None) -> list[dict]:
    """Every combination of the space's values.  With the histogram kernel off
    (``hist_max_bins`` 0) its LDS and workgroup size do not matter, so that
    cutover is tried once."""
    space = dict(DEFAULT_SPACE if space is None else space)
    names = sorted(space)
    out, seen = [], set()
    for vals in itertools.product(*(space[n] for n in names)):
        c = dict(zip(names, (int(v) for v in vals)))
        if c.get("hist_max_bins", 1) == 0:
            c = {k: v for k, v in c.items() if k not in COUNT_KNOBS}
        key = tuple(sorted(c.items()))
        if key not in seen:
            seen.add(key)
            out.append(c)
    return out


def best(trials: list[dict]) -> dict:
    """The trial with the least ms per step (ties: the earlier one)."""
    ok = [t for t in trials if t.get("ms") is not None]
    if not ok:
        raise ValueError("no candidate ran")
    return min(ok, key=lambda t: t["ms"])


def autotune(matrix, group_id, n_targets, opts, space=None, iters=20, warmup=3, repeats=3,
             margin=0.02, log=None) -> dict:
    """Time every candidate of ``space`` on this job in the batched form and
    apply the fastest.  A candidate's ``ms`` is the least, over ``repeats``
    runs, of the mean graph-replay time of ``iters`` replays, after
    ``warmup`` replays (the first candidate gets a second warm-up, so it is
    not timed on a cold device).  The tuning in force stays when it is a
    candidate and the fastest is not faster than it by more than ``margin``
    (a fraction): replay times of near-equal layouts differ by run-to-run
    noise of about that size.  Returns ``{"best": tuning, "trials": [...]}``,
    each trial a candidate with its ``ms`` and its column routing.  A
    candidate the job cannot run under (e.g. a cutover that sends a column
    whose reference slice exceeds the on-chip sort to the sort path) is kept
    with ``ms`` None and its error.  Raises ``TuningMismatch`` when a
    candidate's U1 or p differ from the first runnable candidate's (the
    tuning in force is restored first)."""
    base = current()
    cands = candidates(space)
    if not cands:
        raise ValueError("the tuning space has no candidate")
    trials, ref, timed = [], None, {}
    try:
        for cand in cands:
            # every candidate sorts its own reference: a matrix's kept tables
            # (filled by the first candidate's run) would time later
            # candidates without that sort
            cur = {**base, **cand, "ref_cache": 0}
            try:
                # a tuning the library refuses (e.g. LDS this device does not
                # give) is a candidate that cannot run, not a failed sweep
                apply(cur)
                plan = _capi.Plan(matrix, group_id, n_targets, opts, form=1)
            except _capi.HpdexError as e:
                trials.append({**cand, "ms": None, "error": str(e)})
                if log:
                    log(f"autotune {cand}: not runnable ({e})")
                continue
            info = plan.info()
            # the knobs that shape this job's launches: the histogram knobs
            # only with a histogram column, the sort-path knobs only with a
            # sort-path column (and the routing the cutover gives)
            eff = (info["count_columns"], info["count_blocks"],
                   tuple(int(cur.get(k, 0)) for k in COUNT_KNOBS) if info["count_columns"] else (),
                   tuple(int(cur.get(k, 0)) for k in SORT_KNOBS) if info["split_columns"] else ())
            if eff in timed:
                plan.close()
                t = dict(timed[eff])
                trials.append({**cand, "ms": t["ms"], "count_columns": t["count_columns"],
                               "count_blocks": t["count_blocks"], "same_layout_as": t["index"]})
                if log:
                    log(f"autotune {cand}: same layout as candidate {t['index']}")
                continue
            try:
                # the first candidate that runs gets a second warm-up
                plan.run(warmup * (2 if ref is None else 1))
                U1, P = plan.results()
                ms = min(plan.run(iters)[1] for _ in range(max(1, repeats)))
            except _capi.HpdexError as e:
                trials.append({**cand, "ms": None, "error": str(e)})
                if log:
                    log(f"autotune {cand}: not runnable ({e})")
                continue
            finally:
                plan.close()
            if ref is None:
                ref = (U1.copy(), P.copy())
            elif not (np.array_equal(U1, ref[0]) and np.array_equal(P, ref[1], equal_nan=True)):
                raise TuningMismatch(f"tuning {cand} changed the results")
            timed[eff] = {"ms": ms, "count_columns": info["count_columns"],
                          "count_blocks": info["count_blocks"], "index": len(trials)}
            trials.append({**cand, "ms": ms, "count_columns": info["count_columns"],
                           "count_blocks": info["count_blocks"]})
            if log:
                log(f"autotune {cand}: {ms:.4f} ms")
    except Exception:
        apply(base)
        raise
    if ref is None:
        apply(base)
        raise _capi.HpdexError("no candidate tuning runs this job: " + trials[0]["error"])
    win = best(trials)
    held = [t for t in trials if t["ms"] is not None and
            all(base.get(k) == v for k, v in t.items() if k in KNOBS)]
    if held and win["ms"] >= held[0]["ms"] * (1.0 - margin):
        win = held[0]
    chosen = {**base, **{k: v for k, v in win.items() if k in KNOBS}}
    apply(chosen)
    return {"best": chosen, "ms": win["ms"], "trials": trials}


def save(result: dict, path) -> None:
    with open(path, "w") as f:
        json.dump({"tuning": result["best"], "ms": result.get("ms"),
                   "trials": result.get("trials", [])}, f, indent=1)


def load(path) -> dict:
    """Apply the tuning a ``save`` wrote; returns it."""
    with open(path) as f:
        t = json.load(f)["tuning"]
    apply(t)
    return t

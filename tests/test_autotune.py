"""The autotuner's search logic on the CPU: the candidate set, the choice of
the fastest trial and the JSON round trip (no device calls)."""
import json

import pytest

from hpdex_amd import autotune


def test_candidates_default_space():
    c = autotune.candidates()
    # (2 histogram cutovers x 3 LDS sizes x 3 workgroup sizes, plus the sort
    # path alone, whose LDS and workgroup size do not matter) x 2 split modes
    # x 3 LDS/global sort cutovers
    assert len(c) == (2 * 3 * 3 + 1) * 2 * 3
    assert {"hist_max_bins": 0, "split_arrival": 1, "ref_lds_keys": 8192} in c
    assert {"hist_max_bins": 0, "split_arrival": 0, "ref_lds_keys": 512} in c
    assert len({tuple(sorted(x.items())) for x in c}) == len(c)
    for x in c:
        assert x["split_arrival"] in (0, 1)
        assert x["ref_lds_keys"] in (8192, 2048, 512)
        if x["hist_max_bins"]:
            assert set(x) == {"hist_max_bins", "count_lds_bytes", "count_block", "split_arrival",
                              "ref_lds_keys"}
            assert x["count_block"] in (256, 512, 1024)
            assert x["count_lds_bytes"] % 16 == 0 and 16384 <= x["count_lds_bytes"] <= 155648


def test_candidates_custom_space():
    c = autotune.candidates({"count_block": (256, 512)})
    assert c == [{"count_block": 256}, {"count_block": 512}]


def test_best_picks_least_ms_first_on_ties():
    trials = [{"count_block": 256, "ms": 0.3}, {"count_block": 512, "ms": 0.2},
              {"count_block": 1024, "ms": 0.2}, {"count_block": 128, "ms": None}]
    assert autotune.best(trials)["count_block"] == 512
    with pytest.raises(ValueError):
        autotune.best([{"ms": None}])


def test_save_load_round_trip(tmp_path, monkeypatch):
    applied = []
    monkeypatch.setattr(autotune, "apply", lambda t: applied.append(dict(t)))
    res = {"best": {"hist_max_bins": 2048, "size_ordered": 1, "count_lds_bytes": 53248,
                    "count_block": 512, "count_packed": 1},
           "ms": 0.25, "trials": [{"count_block": 512, "ms": 0.25}]}
    p = tmp_path / "tuning.json"
    autotune.save(res, p)
    assert json.loads(p.read_text())["tuning"] == res["best"]
    assert autotune.load(p) == res["best"]
    assert applied == [res["best"]]


class _FakePlan:
    """Stands in for a device plan: results and timing follow the tuning."""
    tuning = {}
    count_columns = 1
    split_columns = 0
    results_of = staticmethod(lambda t: (None, None))

    def __init__(self, *a, **k):
        self.t = dict(_FakePlan.tuning)

    def run(self, iters=1):
        return 0.0, 1.0 / self.t.get("count_block", 1)

    def results(self):
        return _FakePlan.results_of(self.t)

    def info(self):
        return {"count_columns": _FakePlan.count_columns, "count_blocks": 1,
                "split_columns": _FakePlan.split_columns}

    def close(self):
        pass


@pytest.fixture
def fake_device(monkeypatch):
    import numpy as np
    state = {"count_block": 256}
    monkeypatch.setattr(autotune, "current", lambda: dict(state))
    monkeypatch.setattr(autotune, "apply",
                        lambda t: (state.update(t), setattr(_FakePlan, "tuning", dict(state))))
    monkeypatch.setattr(autotune._capi, "Plan", _FakePlan)
    monkeypatch.setattr(_FakePlan, "count_columns", 1)
    monkeypatch.setattr(_FakePlan, "split_columns", 0)
    same = (np.arange(4.0), np.full(4, 0.5))
    _FakePlan.results_of = staticmethod(lambda t: same)
    return state


def test_autotune_keeps_infeasible_candidates_out(fake_device):
    def results_of(t):
        if t["count_block"] == 1024:
            raise autotune._capi.HpdexError("column 3: too large")
        import numpy as np
        return np.arange(4.0), np.full(4, 0.5)
    _FakePlan.results_of = staticmethod(results_of)
    res = autotune.autotune(None, None, 1, None, space={"count_block": (256, 1024, 512)})
    assert [t["ms"] for t in res["trials"]] == [1 / 256, None, 1 / 512]
    assert res["trials"][1]["error"] == "column 3: too large"
    assert res["best"]["count_block"] == 512 and fake_device["count_block"] == 512


def test_autotune_rejects_a_candidate_that_changes_results(fake_device):
    import numpy as np
    _FakePlan.results_of = staticmethod(
        lambda t: (np.arange(4.0) + (t["count_block"] == 512), np.full(4, 0.5)))
    with pytest.raises(autotune.TuningMismatch):
        autotune.autotune(None, None, 1, None, space={"count_block": (256, 512)})
    assert fake_device["count_block"] == 256      # the tuning in force is restored


def test_autotune_keeps_the_tuning_in_force_within_the_margin(fake_device):
    # 260 is 1.5 % faster than the 256 in force: within the 2 % margin
    res = autotune.autotune(None, None, 1, None, space={"count_block": (256, 260)})
    assert res["best"]["count_block"] == 256 and fake_device["count_block"] == 256
    # with no margin the faster one is taken
    res = autotune.autotune(None, None, 1, None, space={"count_block": (256, 260)}, margin=0.0)
    assert res["best"]["count_block"] == 260


def test_autotune_records_a_tuning_the_library_refuses(fake_device, monkeypatch):
    """set_tuning refusing a candidate (e.g. LDS the device does not give) is
    one candidate that cannot run, recorded with its error; the sweep goes on,
    and the first candidate that runs gets the doubled warm-up."""
    warmups = []
    orig_run = _FakePlan.run

    def run(self, iters=1):
        warmups.append(iters)
        return orig_run(self, iters)

    monkeypatch.setattr(_FakePlan, "run", run)
    state = fake_device

    def apply(t):
        if t.get("count_block") == 1024:
            raise autotune._capi.HpdexError("this device does not give the count kernel its LDS")
        state.update(t)
        _FakePlan.tuning = dict(state)

    monkeypatch.setattr(autotune, "apply", apply)
    res = autotune.autotune(None, None, 1, None, space={"count_block": (1024, 256, 512)},
                            warmup=3, iters=5, repeats=1)
    assert res["trials"][0]["ms"] is None and "LDS" in res["trials"][0]["error"]
    assert [t["ms"] for t in res["trials"][1:]] == [1 / 256, 1 / 512]
    # warm-up of the first runnable candidate doubled (6), the next one's not (3)
    assert warmups == [6, 5, 3, 5]
    assert res["best"]["count_block"] == 512


def test_autotune_empty_space_is_an_error(fake_device):
    with pytest.raises(ValueError, match="no candidate"):
        autotune.autotune(None, None, 1, None, space={"count_block": ()})


def test_autotune_times_one_candidate_per_layout(fake_device, monkeypatch):
    """A job with no histogram column: the histogram knobs do not change its
    launches, so their candidates take the first one's time; the split knob
    does and is timed."""
    monkeypatch.setattr(_FakePlan, "count_columns", 0)
    monkeypatch.setattr(_FakePlan, "split_columns", 5)
    runs = []
    orig = _FakePlan.run

    def run(self, iters=1):
        runs.append(dict(self.t))
        return orig(self, iters)

    monkeypatch.setattr(_FakePlan, "run", run)
    res = autotune.autotune(None, None, 1, None, warmup=1, iters=1, repeats=1,
                            space={"count_block": (256, 512), "split_arrival": (1, 0)})
    timed = [t for t in res["trials"] if "same_layout_as" not in t]
    assert [t["split_arrival"] for t in timed] == [1, 0]
    assert len(res["trials"]) == 4
    for t in res["trials"]:
        if "same_layout_as" in t:
            src = res["trials"][t["same_layout_as"]]
            assert t["ms"] == src["ms"] and t["split_arrival"] == src["split_arrival"]
    assert {r["split_arrival"] for r in runs} == {0, 1}


def test_autotune_sweeps_without_kept_reference_tables(fake_device, monkeypatch):
    """Every candidate sorts its own reference (a matrix's kept tables, filled
    by the first candidate's run, would time the later ones without that
    sort); the tuning applied at the end keeps the caller's ref_cache."""
    fake_device["ref_cache"] = 1
    seen = []
    orig = _FakePlan.__init__

    def init(self, *a, **k):
        orig(self, *a, **k)
        seen.append(self.t.get("ref_cache"))

    monkeypatch.setattr(_FakePlan, "__init__", init)
    res = autotune.autotune(None, None, 1, None, space={"count_block": (256, 512)})
    assert seen == [0, 0]
    assert res["best"]["ref_cache"] == 1 and fake_device["ref_cache"] == 1

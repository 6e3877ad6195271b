"""Autotuner selection logic and tuning-file parsing (no GPU)."""
import json

from idp_amd import autotune, capi


def _row(w, shape, waves, spmv, upd):
    return {"workload": w, "spmv_shape": shape, "upd_waves": waves, "spmv_us": spmv, "update_us": upd}


def test_choose_is_separable_and_sums_workloads():
    rows = []
    # shape 2 wins the sum over two workloads although shape 1 wins the first
    for w, times in (("a", {0: 20.0, 1: 10.0, 2: 11.0, 5: 12.0}), ("b", {0: 20.0, 1: 30.0, 2: 12.0, 5: 13.0})):
        for sh, t in times.items():
            rows.append(_row(w, sh, 4, t, 7.0))
    # update widths measured at shape 2 only
    for w in ("a", "b"):
        rows += [_row(w, 2, 1, 0.0, 9.0), _row(w, 2, 2, 0.0, 6.5), _row(w, 2, 8, 0.0, 6.9)]
        rows += [_row(w, 1, 8, 0.0, 1.0)]  # another shape: ignored for the width
    assert autotune.choose(rows) == (2, 2)


def test_choose_prefers_the_default_on_ties_and_without_data():
    assert autotune.choose([]) == autotune.DEFAULT
    rows = [_row("a", 1, 4, 10.0, 5.0), _row("a", 5, 4, 10.0, 5.0), _row("a", 5, 8, 0.0, 5.0)]
    assert autotune.choose(rows) == (5, 4)


def test_shape_table_matches_the_library_order():
    # lanes per row 4, 8, 16, 32, scalars in wavefront 0 and then in a fifth wavefront
    assert autotune.SHAPE_TABLE[autotune.DEFAULT[0]] == (8, True)
    assert [g for g, _ in autotune.SHAPE_TABLE] == [4, 8, 16, 32] * 2


def test_tuned_config_parsing(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"spmv_shape": 1, "upd_waves": 2}))
    assert capi.tuned_config(str(p)) == (1, 2)
    p.write_text(json.dumps({"spmv_shape": 9, "upd_waves": 2}))
    assert capi.tuned_config(str(p)) is None
    p.write_text("not json")
    assert capi.tuned_config(str(p)) is None
    assert capi.tuned_config(str(tmp_path / "absent.json")) is None


def test_tune_measures_shapes_then_widths(monkeypatch):
    """The search times every shape at the default width, then the other
    widths at the chosen shape, per workload (measurements stubbed)."""
    calls = []

    def fake_measure(kind, sc, shape, waves, newton):
        calls.append((shape, waves))
        spmv = {1: 9.0, 5: 10.0}.get(shape, 20.0)
        upd = {1: 8.0, 2: 6.0, 4: 7.0, 8: 9.0}[waves]
        return {"spmv_shape": shape, "upd_waves": waves, "spmv_us": spmv, "update_us": upd, "samples": 1}

    monkeypatch.setattr(autotune, "measure", fake_measure)
    monkeypatch.setattr(autotune, "_workload", lambda name: ("one", None))
    (shape, waves), rows = autotune.tune(["a", "b"], newton=1, log=lambda *a: None)
    assert (shape, waves) == (1, 2)
    n_shapes = len(autotune.SHAPE_TABLE)
    assert calls[: 2 * n_shapes] == [(s, autotune.DEFAULT[1]) for s in range(n_shapes)] * 2
    assert calls[2 * n_shapes:] == [(1, w) for w in autotune.UPD_WAVES if w != autotune.DEFAULT[1]] * 2
    assert len(rows) == len(calls) and {r["workload"] for r in rows} == {"a", "b"}


def test_choose_team_sums_workloads_and_prefers_the_default():
    rows = [{"workload": "a", "barrier_team": 8, "barrier_hess_us": 30.0},
            {"workload": "a", "barrier_team": 16, "barrier_hess_us": 20.0},
            {"workload": "b", "barrier_team": 8, "barrier_hess_us": 5.0},
            {"workload": "b", "barrier_team": 16, "barrier_hess_us": 20.0}]
    assert autotune.choose_team(rows) == 8
    assert autotune.choose_team([]) == autotune.DEFAULT_TEAM
    tie = [{"workload": "a", "barrier_team": t, "barrier_hess_us": 1.0} for t in autotune.TEAMS]
    assert autotune.choose_team(tie) == autotune.DEFAULT_TEAM


def test_tuned_barrier_team_parsing(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps({"spmv_shape": 1, "upd_waves": 2, "barrier_team": 32}))
    assert capi.tuned_barrier_team(str(p)) == 32
    p.write_text(json.dumps({"barrier_team": 12}))
    assert capi.tuned_barrier_team(str(p)) is None
    assert capi.tuned_barrier_team(str(tmp_path / "absent.json")) is None

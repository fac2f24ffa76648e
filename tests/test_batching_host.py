"""Host logic of the batching runner: tuning-table lookup and stream specs
(no device calls)."""
from libvc1_amd import batching

TABLE = {
    "settings": {"1920x1080": {"jobs_per_step": 255, "schedule": "diagonal", "graphs": True, "mbs_per_s": 7.0e6}},
    "trials": {"1920x1080": {"4r": 3.7e6, "4": 0.8e6, "4g": 0.8e6, "16r": 4.9e6, "16": 2.9e6, "16g": 2.9e6,
                             "64r": 5.4e6, "64": 5.7e6, "64g": 5.8e6, "255r": 5.6e6, "255": 7.2e6, "255g": 7.3e6}},
}


def test_setting_follows_the_pairs_available():
    few = batching.setting_for(TABLE, 1920, 1080, 10)
    assert few["schedule"] == "relax" and not few["graphs"] and few["jobs_per_step"] == 10
    mid = batching.setting_for(TABLE, 1920, 1080, 100)
    assert mid["schedule"] == "diagonal" and mid["graphs"] and mid["jobs_per_step"] == 100
    many = batching.setting_for(TABLE, 1920, 1080, 1000)
    assert many == {"jobs_per_step": 255, "schedule": "diagonal", "graphs": True, "in_flight": 1, "mbs_per_s": 7.3e6}
    # fewer pairs than the smallest trial: that trial's best mode
    assert batching.setting_for(TABLE, 1920, 1080, 2)["schedule"] == "relax"


def test_setting_without_trials_and_legacy_tables():
    assert batching.setting_for(TABLE, 1920, 1080) == {**TABLE["settings"]["1920x1080"], "in_flight": 1}
    legacy = {"settings": {"176x144": {"jobs_per_step": 3, "graphs": True}}}
    assert batching.setting_for(legacy, 176, 144, 5) == {"jobs_per_step": 3, "graphs": True, "schedule": "diagonal",
                                                         "in_flight": 1}
    assert batching.setting_for({}, 640, 360, 7) == batching.DEFAULT_SETTING


def test_setting_with_steps_in_flight():
    table = {"trials": {"1920x1080": {"16rg": 7.9e6, "16rgx2": 9.8e6, "16rgx3": 10.1e6, "64g": 7.0e6}}}
    # enough pairs for three steps of 16 in flight
    s = batching.setting_for(table, 1920, 1080, 48)
    assert s["in_flight"] == 3 and s["schedule"] == "relax" and s["graphs"] and s["jobs_per_step"] == 16
    # 40 pairs fill two steps of 16 but not three
    s = batching.setting_for(table, 1920, 1080, 40)
    assert s["in_flight"] == 2 and s["jobs_per_step"] == 20
    # too few pairs for two steps: one step of the smallest trial size
    s = batching.setting_for(table, 1920, 1080, 20)
    assert s["in_flight"] == 1 and s["jobs_per_step"] == 20


def test_run_in_flight_takes_every_item_once_in_order():
    import threading
    seen, lock = [], threading.Lock()

    def fn(lane, k):
        with lock:
            seen.append((lane, k))
        return k * k
    assert batching.run_in_flight(["a", "b", "c"], 10, fn) == [k * k for k in range(10)]
    assert sorted(k for _, k in seen) == list(range(10))
    assert batching.run_in_flight(["a"], 3, lambda lane, k: (lane, k)) == [("a", 0), ("a", 1), ("a", 2)]


def test_run_in_flight_reraises_the_first_error():
    def fn(lane, k):
        if k == 4:
            raise ValueError("step 4")
        return k
    try:
        batching.run_in_flight(["a", "b"], 8, fn)
    except ValueError as e:
        assert str(e) == "step 4"
    else:
        raise AssertionError("the error of step 4 was not raised")


def test_stream_specs_ranges():
    specs = batching.stream_specs(64, 5)
    assert len(specs) == 64
    assert all((w, h) in batching.RESOLUTIONS and 8 <= n <= 120 for _, w, h, n in specs)
    assert batching.stream_specs(64, 5) == specs        # seeded
    short = batching.stream_specs(64, 5, length_scale=0.01)
    assert all(n == 2 for *_, n in short)

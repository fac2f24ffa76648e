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
    assert many == {"jobs_per_step": 255, "schedule": "diagonal", "graphs": True, "mbs_per_s": 7.3e6}
    # fewer pairs than the smallest trial: that trial's best mode
    assert batching.setting_for(TABLE, 1920, 1080, 2)["schedule"] == "relax"


def test_setting_without_trials_and_legacy_tables():
    assert batching.setting_for(TABLE, 1920, 1080) == dict(TABLE["settings"]["1920x1080"])
    legacy = {"settings": {"176x144": {"jobs_per_step": 3, "graphs": True}}}
    assert batching.setting_for(legacy, 176, 144, 5) == {"jobs_per_step": 3, "graphs": True, "schedule": "diagonal"}
    assert batching.setting_for({}, 640, 360, 7) == batching.DEFAULT_SETTING


def test_stream_specs_ranges():
    specs = batching.stream_specs(64, 5)
    assert len(specs) == 64
    assert all((w, h) in batching.RESOLUTIONS and 8 <= n <= 120 for _, w, h, n in specs)
    assert batching.stream_specs(64, 5) == specs        # seeded
    short = batching.stream_specs(64, 5, length_scale=0.01)
    assert all(n == 2 for *_, n in short)

"""CPU: the committed gfx950 hardware profile (ascend-fft_amd/config/
gfx950_profile.json) is what `python -m bfly.probe` writes from its own probe
record: every field is either a runtime query, the timed copy probe, or one
of the published MI355X figures the probe names -- nothing is hand-set.  The
reference generates its profile the same way (src/host/profile.cpp:1-69,
src/host/bw_probe.cpp:1-53)."""
import json

from bfly import probe

PROFILE = json.loads((probe.PKG / "config" / "gfx950_profile.json").read_text())


def test_profile_is_the_probe_output():
    rec = PROFILE["probe"]
    rebuilt = probe.build_profile(rec["device"], rec["copy"], PROFILE["runtime_version"])
    assert json.loads(json.dumps(rebuilt)) == PROFILE


def test_profile_fields_trace_to_their_sources():
    src = PROFILE["sources"]
    dev = PROFILE["probe"]["device"]
    for key in ("compute_units", "wave_size", "lds_bytes_per_workgroup", "l2_bytes", "global_mem_total_bytes"):
        assert key in src["queried"] and PROFILE[key] == dev[key]
    assert PROFILE["soc"] == dev["gcn_arch"].split(":")[0] == "gfx950"
    assert set(src["published"]) == set(probe.PUBLISHED)
    assert all(PROFILE[k] == v for k, v in probe.PUBLISHED.items())
    copy = PROFILE["probe"]["copy"]
    assert PROFILE["hbm_copy_gbs"] == max(v["gbs"] for v in copy["sizes"].values())
    # a streaming copy reaches most of the nominal peak, never more
    assert 0.5 * PROFILE["hbm_peak_gbs"] < PROFILE["hbm_copy_gbs"] < PROFILE["hbm_peak_gbs"]
    for mib, v in copy["sizes"].items():
        assert abs(2 * int(mib) * 2**20 / (v["us"] * 1e-6) / 1e9 - v["gbs"]) <= 0.1 + 1e-3 * v["gbs"]


def test_mi355x_shape():
    assert PROFILE["compute_units"] == 256 and PROFILE["wave_size"] == 64
    assert PROFILE["lds_bytes_per_cu"] == 160 * 1024
    assert PROFILE["global_mem_total_bytes"] > 250 * 2**30

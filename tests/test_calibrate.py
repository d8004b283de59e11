"""CPU: the plan-selection cost model (ascend-fft_amd/config/cost_model.json)
is the least-squares fit of the committed measured sweep
(config/eta_sweep.json, written on an MI355X by `python -m bfly.calibrate
--measure`): refitting the sweep here gives the committed coefficients, every
kernel the selection can lower to has a fit, and the fit predicts each
measured point within a stated error.  The role of the reference's
scripts/calib_eta.py:1-245 (fit + leave-one-out error)."""
import json

import numpy as np
import pytest

from bfly import calibrate

SWEEP = json.loads(calibrate.SWEEP_FILE.read_text())
MODEL = json.loads(calibrate.MODEL_FILE.read_text())


def test_committed_fit_is_the_refit_of_the_committed_sweep():
    refit = calibrate.fit(SWEEP["points"])
    assert set(refit) == set(MODEL["kernels"])
    for key, m in refit.items():
        assert np.allclose(MODEL["kernels"][key]["c"], m["c"], rtol=1e-6, atol=1e-9), key
        assert MODEL["kernels"][key]["points"] == m["points"]
        assert MODEL["kernels"][key]["loo"] == pytest.approx(m["loo"], abs=1e-6)


def test_every_lowered_kernel_has_a_fit():
    want = {"plain", "stream", "plain_swz", "stream_swz", "stream_lt", "stream2", "split_mfma",
            "split_mfma_stream", "four_step", "plain_single", "stream_lt_single"}
    assert set(MODEL["kernels"]) == want
    assert {calibrate.model_key(p["kernel"], p["n"]) for p in SWEEP["points"]} == want
    assert len(SWEEP["points"]) == len(calibrate.sweep_points())


def test_fit_quality():
    """Leave-one-out mean relative error per kernel, as measured: the model
    ranks kernels, so it needs to be right to tens of percent, not exactly.
    The bound is loose on purpose (launch-bound points vary run to run)."""
    for key, m in MODEL["kernels"].items():
        assert m["loo"] < 0.35, (key, m["loo"])
    refit = calibrate.fit(SWEEP["points"])
    # on the HBM-bound points (>= 2^22 points per launch) the fit is close
    for p in SWEEP["points"]:
        if p["n"] * p["batch"] < (1 << 22):
            continue
        c = refit[calibrate.model_key(p["kernel"], p["n"])]["c"]
        pred = float(np.dot(c, calibrate.features(p["kernel"], p["n"], p["batch"])))
        assert abs(pred - p["us"]) / p["us"] < 0.25, (p, pred)


def test_features_follow_the_chain():
    # the four-step chain moves the tensor twice; one-pass kernels once
    f1 = calibrate.features("stream_lt", 4096, 4096)
    f2 = calibrate.features("four_step", 16384, 1024)
    assert f1[1] == pytest.approx(16 * 4096 * 4096 / 1e6) and f2[1] == pytest.approx(2 * 16 * 16384 * 1024 / 1e6)
    assert calibrate.model_key("plain", 8192) == "plain_single" and calibrate.model_key("plain", 4096) == "plain"
    assert calibrate.model_key("four_step", 8192) == "four_step"

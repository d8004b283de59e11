"""GPU: the predeclared profiles (bfly.suite) and the budgeted soak
(bfly.soak) through bfly_check on the MI355X.

smoke: every case of the smoke profile passes (0 <= maxRel <= 1e-4, exit 0).
long / numeric: the numerical patterns on long chains.  stress: two input
tiers, the 10 000-execution soak with no RSS growth beyond its allowance,
five process lifecycles, and the refused-oversize-allocation probe followed
by a correct transform in the same process."""
import csv
import json

import pytest

from bfly import soak, suite
from conftest import gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X")]


def _need_exe():
    if not suite.CHECK_EXE.exists():
        pytest.fail("bfly_check is not built (make -C ascend-fft_amd)")


@pytest.mark.timeout(400)
@pytest.mark.parametrize("profile,components", [("smoke", "c2c,real,numeric,applications"),
                                                ("long", "numeric")])
def test_profile_passes(tmp_path, profile, components):
    _need_exe()
    rc = suite.main([profile, "--components", components, "--out", str(tmp_path)])
    s = json.loads((tmp_path / "summary.json").read_text())
    bad = [r for r in csv.DictReader((tmp_path / "cases.csv").open()) if r["status"] != "pass"]
    assert rc == 0 and s["result"] == "pass", bad[:5]
    assert s["executions"] == len(suite.cases_of(suite.load_profiles()[profile], components.split(",")))
    assert s["worst_max_rel"] <= 1e-4


@pytest.mark.timeout(400)
def test_soak_lifecycle_and_oom_probe(tmp_path):
    _need_exe()
    rc = soak.main(["--tiers", "1048576,67108864", "--soak-reps", "10000", "--cycles", "5",
                    "--oom-probe", "--device", "0", "--out", str(tmp_path)])
    rows = list(csv.DictReader((tmp_path / "stress.csv").open()))
    bad = [r for r in rows if r["status"] not in ("pass", "refused")]
    assert rc == 0 and not bad, bad[:5]
    assert {r["stage"] for r in rows if r["status"] == "pass"} == {"case", "soak", "lifecycle", "oom_probe"}
    s = json.loads((tmp_path / "summary.json").read_text())
    assert s["result"] == "pass" and s["lifecycle_cycles"] == 5 and len(s["soak_rss_growth_mib"]) == 2

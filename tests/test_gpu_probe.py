"""The random-request ceiling probe used beside the roofline (rmq_probe_gather)."""
import pytest

import rtxrmq_amd as R

pytestmark = pytest.mark.gpu


def test_probe_gather_runs_and_reports_a_rate(gpu):
    r = R.probe_gather(table_bytes=1 << 20, count=1 << 16, reps=2, device=gpu)
    assert r["ms_per_launch"] > 0 and r["gathers_per_s"] > 1e8
    with pytest.raises(ValueError):
        R.probe_gather(table_bytes=3000, device=gpu)

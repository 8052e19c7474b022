"""The serving loop: continuous batching of adaptive sessions of mixed (n, k).

Bars: every admitted session finishes (classified or at the test budget); a session's
experiments, test count and classification are identical whether it is served alone
(per-lattice calls) or in the mix (batched rounds), for both storage forms; admission never
exceeds the HBM budget."""
import pytest

import bayesiangrouptesting_amd as bgt
from bayesiangrouptesting_amd import serving

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [bgt.Precision.float64, bgt.Precision.float32_compensated])
def test_each_session_is_the_same_alone_and_in_the_mix(precision):
    specs = serving.serving_mix(24, rate=0.0, seed=11, lo=0.005, hi=0.10)
    # keep the check quick: sessions up to 2^22 states
    specs = [s for s in specs if s.subjects * s.variants <= 22][:16]
    engine = serving.ServingEngine(precision=precision, hbm_budget=8 << 30)
    report = engine.run(specs, realtime=False)
    assert report["sessions"] == len(specs)
    mixed = {r.sid: r for r in report["results"]}
    for spec in specs:
        alone = serving.serve_alone(spec, precision)
        got = mixed[spec.sid]
        assert (got.experiments, got.tests, got.positive, got.negative) == \
            (alone.experiments, alone.tests, alone.positive, alone.negative), spec.sid


def test_admission_respects_the_hbm_budget():
    specs = serving.serving_mix(30, rate=0.0, seed=3)
    engine = serving.ServingEngine(hbm_budget=1 << 30)
    big = max(engine.footprint(s) for s in specs)
    engine.hbm_budget = max(big, 1 << 30)
    report = engine.run(specs, realtime=False)
    assert report["sessions"] == len(specs)
    assert report["peak_hbm_bytes"] <= engine.hbm_budget
    assert all(r.classified or r.tests == 20 for r in report["results"])

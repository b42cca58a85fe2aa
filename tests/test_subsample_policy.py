"""The auto-subsampling policy (main.rs:66-185 with --enable-auto-subsample),
host arithmetic only: every branch of the reference's decision with its
factor -- coverage from an explicit genome size, else from a HIGH-confidence
estimate; the 1.5x trigger; the clamp to [0.01, 1]; the downgrade to keeping
every read for LOW_COVERAGE_WARNING and UNRELIABLE estimates."""
import pytest

import fastbact_amd as F


def _est(cov=180.0, confidence="HIGH", total_bases=1_134_000_000):
    return F.GenomeEstimate(6_300_000, total_bases // 300, total_bases, 150, 150, cov, 1.0, [77, 99, 127],
                            confidence, 0, 0, 0.0, 0, 0)


@pytest.mark.parametrize("est,depth,g,want", [
    (_est(180.0), 110.0, 0, (110.0 / 180.0, "estimate")),            # config 4: 180x -> 110x
    (_est(165.0), 110.0, 0, (1.0, "within_trigger")),                # not above 1.5 x depth
    (_est(165.0001), 110.0, 0, (110.0 / 165.0001, "estimate")),
    (_est(1e6), 110.0, 0, (0.01, "estimate")),                        # clamp at 0.01
    (_est(500.0, "LOW_COVERAGE_WARNING"), 110.0, 0, (1.0, "estimate_not_high")),
    (_est(500.0, "UNRELIABLE"), 110.0, 0, (1.0, "estimate_not_high")),
    (_est(180.0), 0.0, 0, (1.0, "within_trigger")),                   # depth 0 disables
    (_est(10.0, "UNRELIABLE", total_bases=1_000_000_000), 110.0, 5_000_000, (110.0 / 200.0, "genome_size")),
    (_est(500.0), 110.0, 100_000_000, (1.0, "within_trigger")),       # explicit size: 10x coverage
])
def test_policy_branches(est, depth, g, want):
    f, why = F.auto_subsample_factor(est, depth, 1.5, g)
    assert why == want[1]
    assert f == pytest.approx(want[0], rel=1e-15)


def test_threshold_ratio():
    assert F.auto_subsample_factor(_est(180.0), 110.0, 2.0) == (1.0, "within_trigger")
    f, why = F.auto_subsample_factor(_est(180.0), 110.0, 1.0)
    assert why == "estimate" and f == pytest.approx(110.0 / 180.0)

"""FP64 atomic-rate calibration (cc_atomic_rate), the bound of the voxel tallies.

Every add of the timed and the warm-up launch adds exactly 1.0, so the array total is
exactly the number of adds when none is lost (integers below 2^53 are exact in FP64).
"""
import pytest

pytestmark = pytest.mark.gpu


def test_atomic_rate_counts_every_add(engine):
    n_ops = 4 << 21
    rate, total = engine.atomic_rate(1 << 20, n_ops)
    assert total == 2.0 * n_ops
    assert rate > 1e8


def test_atomic_rate_rejects_bad_sizes(engine):
    for slots, ops in ((0, 1 << 21), (1 << 20, 1000), (1 << 20, 5000 << 21)):
        with pytest.raises(ValueError):
            engine.atomic_rate(slots, ops)

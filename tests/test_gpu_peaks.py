"""GPU online peak detector (chaos_peak_finder) vs the oracle peakFinder:
the hand-traced reference cases plus random quantised signals that exercise
plateaus and the 1e-13 thresholds.  Bit-exact."""
import numpy as np
import pytest

from test_oracle_peaks import HAND_CASES

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("data,peaks,dts", HAND_CASES)
def test_gpu_detector_hand_cases(chaos, data, peaks, dts):
    if len(data) == 0:
        return
    (pk, iv), = chaos.peak_finder(np.array(data, dtype=float), 0.01)
    assert list(pk) == peaks
    assert np.array_equal(iv, np.array(dts, dtype=float) * 0.01)


def test_gpu_detector_random_plateaus(chaos, oracle):
    rng = np.random.default_rng(42)
    rows, n = 512, 400
    steps = rng.choice([0.0, 0.0, 1e-14, -1e-14, 5e-14, -5e-14, 1e-13, 2e-13, -2e-13, 1.0, -1.0, 0.5],
                       size=(rows, n))
    data = np.cumsum(steps, axis=1) + rng.normal(0, 1, (rows, 1))
    print(f"peak_finder rows={rows} n={n}", flush=True)
    got = chaos.peak_finder(data, 0.02)
    for r in range(rows):
        pk, iv = oracle.peak_finder(data[r], 0.02)
        assert np.array_equal(got[r][0], pk), r
        assert np.array_equal(got[r][1], iv), r

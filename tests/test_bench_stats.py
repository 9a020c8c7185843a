"""The bench's latency percentiles (CPU).

The timed steps of bench.py sum 1-us latency histograms instead of keeping
every step's latencies; a percentile is the upper edge of the bin that holds
rank ceil(q n) - 1.  Against the exact order statistic of the same sample it
must never be lower and at most one bin (1 us) higher; an unserved update
(+inf) lands past the histogram, and summing step histograms must give the
percentile of the union.
"""
import numpy as np

import bench


def _exact(lat, q):
    s = np.sort(lat)
    k = min(len(s) - 1, int(np.ceil(q / 100.0 * len(s))) - 1)
    return s[k]


def test_histogram_percentiles_bound_the_order_statistic():
    rng = np.random.default_rng(7)
    for trial in range(30):
        n = int(rng.integers(1, 300000))
        lat = rng.exponential(2.0, n) * rng.choice([1.0, 10.0, 100.0], n)
        if trial % 4 == 0:
            lat[: max(1, n // 500)] = np.inf
        if trial % 5 == 1:
            lat = np.round(lat, 3)   # values on bin edges
        p = bench.hist_percentiles(bench.lat_hist(lat), (50, 99, 99.9))
        for q in (50, 99, 99.9):
            ex = _exact(lat, q)
            if np.isfinite(ex) and ex < bench.LAT_BINS / 1000.0 - 0.001:
                assert ex <= p[q] <= ex + 0.001 + 1e-9, (trial, q, p[q], ex)
            else:
                assert p[q] >= (bench.LAT_BINS - 1) / 1000.0, (trial, q, p[q], ex)


def test_summed_step_histograms_give_the_union():
    rng = np.random.default_rng(11)
    steps = [rng.exponential(1.5, int(rng.integers(1000, 50000))) for _ in range(6)]
    h = sum(bench.lat_hist(x) for x in steps)
    allv = np.concatenate(steps)
    p = bench.hist_percentiles(h, (50, 99, 99.9))
    for q in (50, 99, 99.9):
        ex = _exact(allv, q)
        assert ex <= p[q] <= ex + 0.001 + 1e-9

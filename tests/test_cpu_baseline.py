"""bench.py's CPU baseline leg runs the oracle on a bounded sample (CPU only)."""
import numpy as np

import bench
from oracle import liboracle as lo


def test_cpu_baseline_sample_runs_small():
    o, v = 3, 10
    rng = np.random.default_rng(0)
    x = rng.standard_normal(o * o * v * v) * np.exp(-rng.uniform(0, 6, o * o * v * v))
    m = 300
    idx, val, info = lo.compress(x, m, lo.FRI_SYSTEMATIC, 0.4)
    assert info["status"] == 0
    t, thr, desc = bench.cpu_baseline_sample(x, idx, val, o, v, m, blocks_per_term=4, slice_len=400)
    assert t > 0 and thr >= 1 and "FAITHFUL" in desc

"""The serving and batching workload generators (no GPU)."""
import numpy as np

from lbm_amd.batch import channels_from_seed
from lbm_amd.serve import poisson_requests


def test_poisson_stream_is_seeded_and_in_range():
    a = poisson_requests(200, rate=2.0, seed=7)
    b = poisson_requests(200, rate=2.0, seed=7)
    assert [(r.arrival, r.channel, r.steps, r.storage) for r in a] == \
           [(r.arrival, r.channel, r.steps, r.storage) for r in b]
    arr = np.array([r.arrival for r in a])
    assert np.all(np.diff(arr) > 0)
    gaps = np.diff(np.concatenate([[0.0], arr]))
    assert 0.35 < gaps.mean() < 0.65          # rate 2/s
    for r in a:
        c = r.channel
        assert 256 <= c.lx <= 4096 and c.ly == c.lx // 4
        assert 1000 <= r.steps <= 50000
        assert 20.0 <= c.reynolds <= 500.0
        assert r.storage in ("f64", "f32")
    assert {r.storage for r in a} == {"f64", "f32"}


def test_batch_channels_seeded():
    a = channels_from_seed(64, seed=1)
    b = channels_from_seed(64, seed=1)
    assert a == b
    lxs = np.array([c.lx for c in a])
    assert lxs.min() >= 128 and lxs.max() <= 2048
    # log-uniform: the median sits near the geometric mean of the range
    assert 300 < np.median(lxs) < 900

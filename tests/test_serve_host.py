"""Host logic of the swarm server (no device): arrivals, the convergence rule,
block-mean error estimates and the FIFO latency model of the plain path."""
import numpy as np
import pytest
from scipy import stats

from mcstreamer_amd.serve import (Policy, Request, _host_tally, draw_requests, fifo_latencies,
                                  percentiles, request_dt)
from mcstreamer_amd.swarm import tally_estimates


def test_arrivals_are_poisson_with_the_drawn_ranges():
    reqs = draw_requests(200.0, 30.0, seed=3)
    t = np.array([r.arrival for r in reqs])
    assert abs(len(reqs) - 6000) < 5 * np.sqrt(6000)
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert stats.kstest(gaps * 200.0, "expon").pvalue > 1e-3
    en = np.array([r.en_td for r in reqs])
    assert en.min() >= 10 and en.max() <= 1000
    assert stats.kstest((np.log10(en) - 1) / 2, "uniform").pvalue > 1e-3
    n0 = np.array([r.n0 for r in reqs])
    assert n0.min() >= 1000 and n0.max() <= 1_000_000
    tg = np.array([r.target for r in reqs])
    assert tg.min() >= 0.005 and tg.max() <= 0.02
    assert request_dt(100.0) == pytest.approx(1e-12) and request_dt(1000.0) == pytest.approx(1e-13)


def test_host_tally_block_means_and_error():
    rng = np.random.default_rng(1)
    relax, bs = 10, 5
    means = rng.normal(size=(10 + 5 * 8 + 3, 4)) + 10.0
    t = _host_tally(means, relax, bs)
    assert t.nrecorded == 43 and t.nblocks == 8
    est = tally_estimates(t)
    rec = means[relax:]
    bm = rec[:40].reshape(8, 5, 4).mean(axis=1)
    for q, name in enumerate(("drift_velocity", "mean_energy", "ionization_rate", "attachment_rate")):
        assert est[name][0] == pytest.approx(rec[:, q].mean(), rel=1e-13)
        se = bm[:, q].std(ddof=1) / np.sqrt(8) / abs(bm[:, q].mean())
        assert est[name][1] == pytest.approx(se, rel=1e-9)


def test_convergence_rule():
    p = Policy(relax_steps=100, block_steps=10, min_blocks=5, max_steps=1000)
    assert not p.done(149, 0.0, 0.01)
    assert p.done(150, 0.005, 0.01)
    assert not p.done(150, 0.02, 0.01)
    assert p.done(1000, 1.0, 0.01)


def test_fifo_latency_model():
    reqs = [Request(k, t, 100.0, 0.0, 1000, 0.01) for k, t in enumerate((0.0, 0.5, 0.6, 5.0))]
    lat = fifo_latencies(reqs, [1.0, 1.0, 1.0, 1.0])
    np.testing.assert_allclose(lat, [1.0, 1.5, 2.4, 1.0])
    assert percentiles(lat)["p50"] == pytest.approx(1.25)


def test_checkpoint_rejects_foreign_files(tmp_path):
    """A checkpoint file of another format is refused before anything is built."""
    import json

    import numpy as np
    import pytest

    from mcstreamer_amd import checkpoint
    path = str(tmp_path / "other.npz")
    np.savez(path, meta=np.array(json.dumps({"format": "something else"})))
    with pytest.raises(ValueError):
        checkpoint.load(path)

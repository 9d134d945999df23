"""Host side of the swarm statistics (Swarm/src/Swarm.jl:253-306), no GPU:

- ks_statistic: the reference's two-sample statistic (HypothesisTests' ksstats,
  called at Swarm.jl:297) against scipy's ks_2samp on tie-free samples, and its
  documented behaviour on ties (it evaluates the running difference inside a tie).
- kolmogorov_sf: the two published series of the Kolmogorov distribution against
  scipy's kstwobign on both branches (t <= 1 and t > 1).
- ks_pvalue: P(K > sqrt(n m / (n + m)) delta), the approximate test's p-value.
- stats: mean and std/sqrt(n) (n - 1 divisor) per column, as Swarm.stats.

The p-values of HypothesisTests itself are not importable here: parity with the
reference's exact p-values is unpinned; these tests pin the published algorithm."""
import numpy as np
import pytest
import scipy.stats as ss

from mcstreamer_amd import swarm as S


@pytest.mark.parametrize("nx,ny,shift", [(500, 700, 0.1), (64, 64, 0.0), (1000, 37, 0.5), (3, 5, 1.0)])
def test_ks_statistic_matches_scipy_without_ties(nx, ny, shift):
    rng = np.random.default_rng(nx + ny)
    x = rng.normal(size=nx)
    y = rng.normal(shift, 1.0, size=ny)
    d, dp, dn = S.ks_statistic(x, y)
    r = ss.ks_2samp(x, y)
    assert d == pytest.approx(r.statistic, abs=1e-14)
    assert d == max(dp, dn)
    # one-sided parts: sup(F_x - F_y) and sup(F_y - F_x)
    assert dp == pytest.approx(ss.ks_2samp(x, y, alternative="greater").statistic, abs=1e-14)
    assert dn == pytest.approx(ss.ks_2samp(x, y, alternative="less").statistic, abs=1e-14)


def test_ks_statistic_inside_ties_like_the_reference():
    # pooled [1, 1, 1] in stable order x, x, y: running sums 1/2, 1, 0
    d, dp, dn = S.ks_statistic([1.0, 1.0], [1.0])
    assert (d, dp, dn) == (1.0, 1.0, 0.0)
    # without ties the same samples shifted apart give the same supremum
    assert S.ks_statistic([1.0, 1.0], [2.0])[0] == 1.0


@pytest.mark.parametrize("t", [0.05, 0.2, 0.5, 0.8, 0.999, 1.0, 1.001, 1.3, 2.0, 3.0, 5.0])
def test_kolmogorov_sf_matches_scipy(t):
    assert S.kolmogorov_sf(t) == pytest.approx(ss.kstwobign.sf(t), rel=1e-12, abs=1e-300)


def test_kolmogorov_sf_edges():
    assert S.kolmogorov_sf(0.0) == 1.0 and S.kolmogorov_sf(-1.0) == 1.0
    assert S.kolmogorov_sf(10.0) == 0.0 or S.kolmogorov_sf(10.0) < 1e-80
    # monotone decreasing
    ts = np.linspace(0.05, 4, 200)
    v = np.array([S.kolmogorov_sf(t) for t in ts])
    assert np.all(np.diff(v) <= 1e-15)


def test_ks_pvalue_is_the_limit_law_at_the_scaled_statistic():
    rng = np.random.default_rng(3)
    x, y = rng.exponential(size=800), rng.exponential(1.05, size=600)
    d = ss.ks_2samp(x, y).statistic
    want = ss.kstwobign.sf(np.sqrt(800 * 600 / 1400) * d)
    assert S.ks_pvalue(x, y) == pytest.approx(want, rel=1e-10)
    # same distribution, large samples: p is not small; shifted: p is tiny
    a, b = rng.normal(size=20000), rng.normal(size=20000)
    assert S.ks_pvalue(a, b) > 1e-3
    assert S.ks_pvalue(a, b + 0.2) < 1e-20


def test_stats_mean_and_standard_error():
    rng = np.random.default_rng(5)
    dfs = [dict(en=np.array([1.0, 10.0]), mobility=rng.normal(size=2), k=rng.normal(size=(2, 3)))
           for _ in range(7)]
    st = S.stats(dfs)
    mob = np.array([d["mobility"] for d in dfs])
    np.testing.assert_allclose(st["mobility"], mob.mean(axis=0), rtol=1e-15)
    np.testing.assert_allclose(st["mobility_err"], mob.std(axis=0, ddof=1) / np.sqrt(7), rtol=1e-14)
    ks = np.array([d["k"] for d in dfs])
    np.testing.assert_allclose(st["k_err"], ks.std(axis=0, ddof=1) / np.sqrt(7), rtol=1e-14)
    assert "en_err" not in st and np.array_equal(st["en"], [1.0, 10.0])
    with pytest.raises(ValueError):
        S.stats([dict(en=1.0, a=1.0), dict(en=1.0, b=1.0)])

"""Pins of the random stream and the elementary functions the step uses.

Philox4x32-10 against the published known-answer vectors of Salmon et al.
(Random123, SC'11, kat_vectors); the stream's word-to-double mapping by its
definition; log/exp/sincos against numpy's libm; the Poisson, samplemin,
nextcoll and randsphere draws against their distributions."""
import numpy as np
import pytest

from oracle import refcpu as R

KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_kat(ctr, key, out):
    assert R.philox(ctr, key) == out


def test_stream_mapping_matches_definition():
    seed, inst, phase, pid, step = 0x1234_5678_9ABC_DEF0, 3, 1, (5 << 32) + 77, 42
    u = R.stream_u01(seed, inst, phase, pid, step, 6)
    key = [seed & 0xFFFFFFFF, ((seed >> 32) + inst) & 0xFFFFFFFF]
    exp = []
    for blk in range(3):
        w = R.philox([blk, step, pid & 0xFFFFFFFF, ((pid >> 32) & 0xFFFFFF) | (phase << 24)], key)
        for a, b in ((w[0], w[1]), (w[2], w[3])):
            exp.append(((a >> 5) * 67108864.0 + (b >> 6)) / 9007199254740992.0)
    np.testing.assert_array_equal(u, exp)


def test_streams_uniform_and_distinct():
    u = R.draws(0, 200_000)
    assert 0.0 <= u.min() and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5 * np.sqrt(1 / 12 / u.size)
    hist, _ = np.histogram(u, bins=50, range=(0, 1))
    chi2 = ((hist - u.size / 50) ** 2 / (u.size / 50)).sum()
    assert chi2 < 100  # 49 dof
    a = R.stream_u01(1, 0, 0, 9, 0, 8)
    b = R.stream_u01(1, 0, 1, 9, 0, 8)
    c = R.stream_u01(1, 1, 0, 9, 0, 8)
    d = R.stream_u01(1, 0, 0, 9, 1, 8)
    assert len({tuple(a), tuple(b), tuple(c), tuple(d)}) == 4


def test_log_exp_sincos_against_libm():
    rng = np.random.default_rng(0)
    x = np.concatenate([np.exp(rng.uniform(-700, 700, 50_000)), rng.uniform(1e-12, 1.0, 50_000),
                        rng.uniform(0.5, 2.0, 50_000), [1.0, 2.0, 0.5, np.sqrt(2.0)]])
    lg, _ = R.math(0, x)
    ref = np.log(x)
    ulp = np.abs(lg - ref) / np.spacing(np.maximum(np.abs(ref), 1e-300))
    assert np.max(ulp[np.abs(ref) > 1e-3]) <= 4
    assert lg[-4] == 0.0
    y = rng.uniform(-700, 700, 50_000)
    ex, _ = R.math(1, y)
    np.testing.assert_allclose(ex, np.exp(y), rtol=5e-16, atol=0)
    u = rng.uniform(0, 1, 50_000)
    s, c = R.math(2, u)
    np.testing.assert_allclose(s, np.sin(2 * np.pi * u), rtol=0, atol=2e-15)
    np.testing.assert_allclose(c, np.cos(2 * np.pi * u), rtol=0, atol=2e-15)
    s, c = R.math(2, np.array([0.0, 0.25, 0.5, 0.75]))
    np.testing.assert_array_equal(s, [0.0, 1.0, 0.0, -1.0])
    np.testing.assert_array_equal(c, [1.0, 0.0, -1.0, 0.0])


@pytest.mark.parametrize("lam", [0.05, 0.9, 7.0, 820.0])
def test_poisson_sampler_moments(lam):
    n = 100_000 if lam < 100 else 20_000
    k = R.draws(2, n, lam)
    assert np.all(k == np.round(k)) and k.min() >= 0
    se = np.sqrt(lam / n)
    assert abs(k.mean() - lam) < 5 * se
    assert abs(k.var() / lam - 1.0) < 0.05


def test_nextcoll_samplemin_randsphere():
    e = R.draws(1, 200_000)
    assert abs(e.mean() - 1.0) < 0.01 and abs(e.var() - 1.0) < 0.02 and np.isfinite(e).all()
    for m in (1.0, 8.0, 64.0):
        s = R.draws(3, 100_000, m)
        assert abs(s.mean() - 1 / (m + 1)) < 5 * np.sqrt(m / ((m + 1) ** 2 * (m + 2)) / s.size)
    o = R.draws(4, 100_000)
    np.testing.assert_allclose(np.linalg.norm(o, axis=1), 1.0, atol=1e-15)
    assert np.all(np.abs(o.mean(axis=0)) < 0.01)
    np.testing.assert_allclose((o ** 2).mean(axis=0), 1 / 3, atol=0.01)

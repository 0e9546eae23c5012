"""The SK threshold restatement (SKLimits.C:136-229) pinned by what fixes it
independently: the Pearson density it builds integrates to one with the mean
and variance it was built from, the thresholds bound the right tail mass of a
simulated SK estimator, and large counts follow the closed form."""
import importlib
import math

import numpy as np
import pytest

S = importlib.import_module("dspsr-kanyixia_amd.sklimits")


@pytest.mark.parametrize("M,Nd", [(16, 1), (64, 1), (1024, 1), (128, 4)])
def test_pearson_density_moments(M, Nd):
    from scipy import integrate
    L = S.SKLimits(M, 3.0, Nd, 1)
    if Nd == 1:
        _, mu2, _, _ = S._moments_iv(float(M))
    else:
        mu2 = S._moments_vi(float(M), float(Nd))[0]
    sd = math.sqrt(mu2)
    pts = [1 - 2 * sd, 1, 1 + 2 * sd]
    lo, hi = 1 - 60 * sd, 1 + 60 * sd
    norm = integrate.quad(L.pdf, lo, hi, points=pts, limit=500, epsrel=1e-12)[0]
    mean = integrate.quad(lambda x: x * L.pdf(x), lo, hi, points=pts, limit=500, epsrel=1e-12)[0]
    var = integrate.quad(lambda x: (x - 1) ** 2 * L.pdf(x), lo, hi, points=pts, limit=500, epsrel=1e-12)[0]
    assert abs(norm - 1) < 1e-6
    assert abs(mean - 1) < 1e-6
    assert abs(var / mu2 - 1) < 1e-4


def _shape_moments(M, Nd):
    """(mu2, beta1, beta2) of SKLimits: parametersIV (Nd = 1) or the
    parametersVI moments (SKLimits.C:70-134)."""
    if Nd == 1:
        _, mu2, b1, b2 = S._moments_iv(float(M))
    else:
        mu2, mu3, mu4 = S._moments_vi(float(M), float(Nd))
        b1, b2 = mu3 ** 2 / mu2 ** 3, mu4 / mu2 ** 2
    return mu2, b1, b2


@pytest.mark.parametrize("M,Nd", [(64, 1), (1024, 1), (128, 4), (16, 1), (1024, 16)])
def test_pearson_density_shape(M, Nd):
    """The third and fourth central moments of the density SKLimits builds:
    a Pearson IV curve (0 < kappa < 1) carries all four moments it was
    built from, so its beta1 = mu3^2/mu2^3 and beta2 = mu4/mu2^2 equal
    the parameters; a Pearson VI curve (kappa > 1: M = 16, and M = 1024 at
    Nd = 16) is fixed by mu2 and mu3, so its beta1 equals the parameter
    and its beta2 is the family's own."""
    from scipy import integrate
    L = S.SKLimits(M, 3.0, Nd, 1)
    mu2, b1, b2 = _shape_moments(M, Nd)
    sd = math.sqrt(mu2)
    pts = [1 - 2 * sd, 1, 1 + 2 * sd]
    lo, hi = max(1 - 60 * sd, -50.0), 1 + 200 * sd

    def mom(k):
        return integrate.quad(lambda x: (x - 1) ** k * L.pdf(x), lo, hi, points=pts, limit=1000, epsrel=1e-12)[0]

    m2, m3, m4 = mom(2), mom(3), mom(4)
    assert abs(m3 ** 2 / m2 ** 3 / b1 - 1) < 1e-4
    if isinstance(L.pdf, S._PearsonIV):
        assert abs(m4 / m2 ** 2 / b2 - 1) < 1e-4
    else:
        assert (M, Nd) in ((16, 1), (1024, 16))


@pytest.mark.parametrize("M,Nd", [(64, 1), (1024, 1), (128, 4)])
def test_sk_moments_monte_carlo(M, Nd):
    """The moments SKLimits fits (mu2, beta1, beta2 of Nita & Gary; its
    mean is 1) against those of the SK estimator itself, 10^6 trials of
    M samples of power distributed as the sum of Nd unit exponentials
    (|x|^2 of complex Gaussian noise, integrated Nd times), SK =
    (M Nd + 1)/(M - 1) (M S2 / S1^2 - 1).  Each sample moment agrees within
    five of its standard errors (from 100 batches of 10^4 trials)."""
    rng = np.random.default_rng(M + Nd)
    nb, per = 100, 10_000
    est = np.empty((nb, 4))
    for b in range(nb):
        p = rng.standard_exponential((per, M)) if Nd == 1 else rng.standard_gamma(Nd, (per, M))
        s1 = p.sum(1)
        s2 = np.einsum("ij,ij->i", p, p)
        sk = (M * Nd + 1) / (M - 1) * (M * s2 / (s1 * s1) - 1)
        d = sk - sk.mean()
        v = np.mean(d * d)
        est[b] = sk.mean(), v, np.mean(d ** 3) ** 2 / v ** 3, np.mean(d ** 4) / v ** 2
    mean, se = est.mean(0), est.std(0, ddof=1) / math.sqrt(nb)
    mu2, b1, b2 = _shape_moments(M, Nd)
    for got, want, e in zip(mean, (1.0, mu2, b1, b2), se):
        assert abs(got - want) < 5 * e, (got, want, e)
    # and resolves a plausible slip: mu2 with (M + 1) for (M + 2) in its
    # denominator is 1.7% off at M = 64, eight standard errors
    if (M, Nd) == (64, 1):
        slip = 4.0 * M * M / ((M - 1) * (M + 1) * (M + 3))
        assert abs(mean[1] - slip) > 5 * se[1]


def test_thresholds_bound_simulated_tail_mass():
    """SK of complex Gaussian noise (|x|^2 exponential) over M = 64 samples,
    2e5 trials: each side's exceedance is near the per-comparison target.
    The Pearson IV curve is the reference's approximation to the estimator's
    distribution: at M = 64 its lower tail is ~40% light (0.0008 against
    0.00135, which 4e5 trials resolve), so the bound is a factor of 1.5."""
    M, trials = 64, 200_000
    L = S.SKLimits(M, 3.0, 1, 1)
    rng = np.random.default_rng(5)
    p = rng.exponential(size=(trials, M))
    s1, s2 = p.sum(1), (p * p).sum(1)
    sk = (M + 1) / (M - 1) * (M * s2 / (s1 * s1) - 1)
    target = (1 - math.erf(3 / math.sqrt(2))) / 2 * trials  # 270
    for n in (np.sum(sk > L.upper), np.sum(sk < L.lower)):
        assert 0.5 * target < n < 1.5 * target, (n, target)
    # the symmetric Gaussian bounds are visibly wrong on the upper side
    assert np.sum(sk > 1 + L.symmetric) > 1.5 * target


def test_large_count_closed_form():
    L = S.SKLimits(40000, 3.0, 1, 2)
    pcer = 1 - (1 - (1 - math.erf(3 / math.sqrt(2)))) ** 0.5
    pcsd = math.sqrt(2) * __import__("scipy").special.erfinv(1 - pcer)
    f = math.sqrt(4 / 40000) * pcsd
    assert abs(L.lower - (1 - f)) < 1e-12 and abs(L.upper - (1 + f)) < 1e-12


def test_limit_tables_subset():
    lo, up, sym, pcsd = S.limit_tables(64, 3.0, counts=[64, 2])
    assert lo[64] > 0 and up[64] > 1 and lo[2] == 0 and lo[63] == 0
    L = S.SKLimits(64, 3.0, 1, 2)
    assert lo[64] == np.float32(L.lower) and up[64] == np.float32(L.upper)


def test_committed_tables_match_the_restatement():
    """tests/golden/sk_limits.npz (used by the GPU tests) is what
    limit_tables gives (make_sk_limits.py wrote it)."""
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sk_limits.npz"))
    lo, up, sym, pcsd = S.limit_tables(64, 3.0, npol=2, counts=[64])
    assert np.array_equal(z["lower64"], lo) and np.array_equal(z["upper64"], up) and z["pcsd64"] == pcsd
    lo, up, sym, pcsd = S.limit_tables(256, 3.0, npol=2, counts=[256])
    assert np.array_equal(z["lower256"], lo) and np.array_equal(z["upper256"], up) and z["pcsd256"] == pcsd
    assert np.array_equal(z["sym256"], sym)
    lo, up, sym, pcsd = S.limit_tables(32, 3.0, npol=2, counts=[3, 6, 17, 32])
    for c in (3, 6, 17, 32):
        assert z["lower32"][c] == lo[c] and z["upper32"][c] == up[c] and z["sym32"][c] == sym[c]

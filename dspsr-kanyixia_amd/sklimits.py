"""Spectral-kurtosis thresholds by count: the host side of the SK detector.

dspsr's SpectralKurtosis transformation evaluates its thresholds on the host
(SpectralKurtosis::get_limits, SpectralKurtosis.C:989-1009) and hands them to
its engine; the HIP engine here takes them as tables indexed by the block
count M' (dspx_sk_set_limits).  This module computes those tables for callers
that are not dspsr itself (the tests, bench.py, the dspx_pipeline command).

It restates SKLimits::calc_limits (Signal/Statistics/SKLimits.C:136-229): the
SK estimator's distribution is approximated by a Pearson type IV density
(PearsonIV.C, moments of SKLimits.C:70-89) or, when kappa > 1, type VI
(PearsonVI.C, SKLimits.C:113-134); the cumulative and complementary
cumulative functions are integrated numerically and inverted with the same
clamped Newton-Raphson iteration (dsp/NewtonRaphson.h: at most 100 steps,
relative precision 1e-12, limits 1 -/+ 2*symmetric, the initial guess kept
when the iteration fails).  The reference integrates with a Romberg
midpoint rule (dsp/Romberg.h); this uses scipy's adaptive quadrature, so the
thresholds agree to the integration tolerance, not bit for bit.
"""
import math

import numpy as np

__all__ = ["SKLimits", "limit_tables"]


def _loggammar2(x, y):
    """log |Gamma(x+iy) / Gamma(x)|^2 by the series of Abramowitz & Stegun
    (eq. 6.1.25 summed as in Heinrich's Pearson IV guide)."""
    y2 = y * y
    xmin = max(2 * y2, 10.0)
    r = 0.0
    while x < xmin:
        t = y / x
        x += 1
        r += math.log(1 + t * t)
    s, p, f = 1.0, 1.0, 0.0
    while p > s * 2.220446049250313e-16:
        p *= y2 + f * f
        f += 1
        p /= x * f
        x += 1
        s += p
    return -r - math.log(s)


class _PearsonIV:
    def __init__(self, mu1, mu2, beta1, beta2):
        r = 6 * (beta2 - beta1 - 1) / (2 * beta2 - 3 * beta1 - 6)
        self.m = (r + 2) / 2
        self.nu = -(r * (r - 2) * math.sqrt(beta1)) / math.sqrt(16 * (r - 1) - beta1 * (r - 2) ** 2)
        self.a = 0.25 * math.sqrt(mu2 * (16 * (r - 1) - beta1 * (r - 2) ** 2))
        self.lam = mu1 - 0.25 * (r - 2) * math.sqrt(mu2) * math.sqrt(beta1)
        if self.m <= 0.5:
            raise ValueError("Pearson IV: m <= 0.5")
        log_beta = math.lgamma(self.m - 0.5) + 0.5 * math.log(math.pi) - math.lgamma(self.m)
        self.logk = _loggammar2(self.m, 0.5 * self.nu) - math.log(self.a) - log_beta

    def __call__(self, x):
        a1 = (x - self.lam) / self.a
        return math.exp(self.logk - self.m * math.log1p(a1 * a1) - self.nu * math.atan(a1))


class _PearsonVI:
    def __init__(self, mu2, mu3):
        sq = math.sqrt(16 * mu2 ** 4 + (4 * mu2 + 1) * mu3 * mu3)
        self.alpha = mu3 ** -3 * (32 * mu2 ** 5 - 4 * mu3 * mu2 ** 3 + 8 * mu3 ** 2 * mu2 ** 2 + mu3 ** 2 * mu2
                                  - mu3 ** 3 + (8 * mu2 ** 3 - mu3 * mu2 + mu3 ** 2) * sq)
        self.beta = 3 + 2 * mu2 * mu3 ** -2 * (4 * mu2 ** 2 + sq)
        self.delta = (self.beta - self.alpha - 1) / (self.beta - 1)
        self.logk = math.lgamma(self.alpha + self.beta) - math.lgamma(self.alpha) - math.lgamma(self.beta)

    def __call__(self, x):
        x -= self.delta
        if x <= 0:
            return 0.0
        return math.exp(self.logk + (self.alpha - 1) * math.log(x) - (self.alpha + self.beta) * math.log1p(x))


def _moments_iv(M):
    """SKLimits::parametersIV(M), SKLimits.C:70-89 (Nd = 1)."""
    mu2 = (4 * M * M) / ((M - 1) * (M + 2) * (M + 3))
    beta1 = 4 * (M + 2) * (M + 3) * (5 * M - 7) ** 2 / ((M - 1) * (M + 4) ** 2 * (M + 5) ** 2)
    beta2 = 3 * (M + 2) * (M + 3) * (M ** 3 + 98 * M * M - 185 * M + 78) / ((M - 1) * (M + 4) * (M + 5) * (M + 6) * (M + 7))
    return 1.0, mu2, beta1, beta2


def _moments_vi(M, Nd):
    """SKLimits::parametersVI(M, Nd), SKLimits.C:113-134: mu2, mu3, mu4."""
    b = M / (M - 1)
    lg = math.lgamma(M * Nd + 2)
    mu2 = 2 * Nd * (Nd + 1) * M * b * math.exp(lg - math.lgamma(M * Nd + 4))
    mu3 = 8 * Nd * (Nd + 1) * M * b ** 2 * math.exp(lg - math.lgamma(M * Nd + 6))
    mu3 *= (Nd + 4) * M * Nd - 5 * Nd - 2
    mu4 = 12 * Nd * (Nd + 1) * M * b ** 3 * math.exp(lg - math.lgamma(M * Nd + 8))
    mu4 *= ((M ** 3 + 3 * M ** 2) * Nd ** 4 + (M ** 3 + 68 * M ** 2 - 93 * M) * Nd ** 3
            + (125 * M ** 2 - 245 * M + 84) * Nd ** 2 + (48 - 32 * M) * Nd + 24)
    return mu2, mu3, mu4


class SKLimits:
    """SKLimits(M, std_devs) with set_Nd / set_ntest, calc_limits() run on
    construction: lower, upper, symmetric thresholds (doubles) and the
    per-comparison sigma pcsd."""

    def __init__(self, M, std_devs, Nd=1, ntest=1):
        from scipy import integrate, special
        self.M, self.Nd, self.ntest = int(M), int(Nd), int(ntest)
        if self.M < 2:
            raise ValueError("SKLimits: M=%d < 2" % self.M)
        # SKLimits.C:48: the sigma is rounded to float before erf
        fwer = 1.0 - math.erf(float(np.float32(std_devs)) / math.sqrt(2))
        pcer = 1.0 - (1 - fwer) ** (1.0 / self.ntest)
        target = pcer / 2.0
        self.pcsd = float(special.erfinv(1.0 - pcer)) * math.sqrt(2)
        self.symmetric = 0.0
        M = float(self.M)
        if self.M >= 32768:
            factor = math.sqrt(4.0 / M) * self.pcsd
            self.lower = 1.0 - factor
            self.upper = 1.0 + factor
            return
        if self.Nd == 1:
            mu1, mu2, beta1, beta2 = _moments_iv(M)
        else:
            mu2, mu3, mu4 = _moments_vi(M, float(self.Nd))
            mu1, beta1, beta2 = 1.0, mu3 ** 2 / mu2 ** 3, mu4 / mu2 ** 2
        kappa = beta1 * (beta2 + 3) ** 2 / (4 * (4 * beta2 - 3 * beta1) * (2 * beta2 - 3 * beta1 - 6))
        if kappa < 0:
            raise ValueError("SKLimits: Pearson type I (kappa < 0) not implemented")
        if kappa > 1:
            self.pdf = _PearsonVI(*_moments_vi(M, float(self.Nd))[:2])
        else:
            self.pdf = _PearsonIV(mu1, mu2, beta1, beta2)
        sd = math.sqrt(mu2)
        self.symmetric = sd * self.pcsd
        pdf = self.pdf
        # split the improper integrals far out in the tails, where the density
        # is negligible next to the targets
        far = 40 * sd

        def cf(x):
            a = integrate.quad(pdf, -np.inf, 1 - far, epsabs=0, epsrel=1e-12, limit=400)[0]
            return a + integrate.quad(pdf, 1 - far, x, epsabs=0, epsrel=1e-12, limit=400)[0]

        def ccf(x):
            a = integrate.quad(pdf, 1 + far, np.inf, epsabs=0, epsrel=1e-12, limit=400)[0]
            return a + integrate.quad(pdf, x, 1 + far, epsabs=0, epsrel=1e-12, limit=400)[0]

        lo_lim, hi_lim = 1 - 2 * self.symmetric, 1 + 2 * self.symmetric

        def newton(f, df, y, x):
            for _ in range(100):
                dx = (f(x) - y) / df(x)
                x -= dx
                x = min(max(x, lo_lim), hi_lim)
                if dx * dx <= (x * 1e-12) ** 2:
                    return x
            raise ArithmeticError("Newton-Raphson did not converge")

        y = math.log(target)
        guess = 1 - self.symmetric
        try:
            self.lower = newton(lambda x: math.log(cf(x)), lambda x: pdf(x) / cf(x), y, guess)
        except (ArithmeticError, ValueError, ZeroDivisionError):
            self.lower = guess
        guess = 1 + self.symmetric
        try:
            self.upper = newton(lambda x: math.log(ccf(x)), lambda x: -pdf(x) / ccf(x), y, guess)
        except (ArithmeticError, ValueError, ZeroDivisionError):
            self.upper = guess


def limit_tables(M, std_devs, Nd=1, npol=2, counts=None):
    """Threshold tables indexed by count 0..M, as SpectralKurtosis::get_limits
    builds them (ntest = npol of the sums): float32 lower, upper, symmetric
    and the per-comparison sigma for the closed form of large counts.
    Counts below 3 are never tested by detect_skfb; their entries are 0.
    Counts for which SKLimits::calc_limits throws (Pearson type I) get
    -inf / +inf, so that they are never zapped.
    ``counts``: evaluate only these counts (each takes ~0.1 s; input without
    zeroed samples has every block at count M), leaving the rest 0."""
    lower = np.zeros(M + 1, np.float32)
    upper = np.zeros(M + 1, np.float32)
    sym = np.zeros(M + 1, np.float32)
    pcsd = None
    for c in (range(3, M + 1) if counts is None else sorted(set(int(c) for c in counts if 3 <= c <= M))):
        try:
            L = SKLimits(c, std_devs, Nd, npol)
        except ValueError:
            # counts whose estimator would need a Pearson type I curve (3..5
            # at Nd = 1), where the reference's calc_limits throws: never zapped
            lower[c], upper[c], sym[c] = -np.inf, np.inf, np.inf
            continue
        lower[c], upper[c], sym[c] = L.lower, L.upper, L.symmetric
        pcsd = L.pcsd
    if pcsd is None:
        pcsd = SKLimits(max(M, 2), std_devs, Nd, npol).pcsd
    return lower, upper, sym, pcsd

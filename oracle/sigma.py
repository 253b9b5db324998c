"""Independent restatement of the reference's scattering model, for the oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and through oracle.py by the
smoke check and the CPU baseline of bench.py), never by the product package.

It restates, in numpy/scipy and without reading the product's material code:
  - the PSDF laws and their correlation functions (MaterialClass.m:479-710),
    with the Fourier convention stated at MaterialClass.m:459-477:
    Phi(k) = (2 pi)^-d int exp(-i k.x) R(|x|) dx;
  - the isotropic differential cross-sections sigma(theta) of
    calcSigmaIsotropic (MaterialClass.m:200-357): acoustic 2-D/3-D, elastic
    2-D P-SV and 3-D, four channels P->P, P->S, S->P, S->S;
  - prepareSigma / prepareSigmaOne (MaterialClass.m:2379-2463): Sigma and
    Sigma' by quadrature, mean free times, P2P / S2S, and the inverse-CDF
    knots (1e6-point angle grid, cumulative sum, knots where the cdf step
    exceeds 1e-8 together with their right neighbours).

Sigma here is integrated by composite Gauss-Legendre quadrature with panel
breaks at the kinks of the PSDF (a different integrator from the product's),
so that the product's QUADPACK integration is checked, not repeated.

What pins this file (tests/test_sigma_pins.py): the transform pair Phi <-> R
and the unit variance R(0) = int Phi d^dk = 1 of every law, the correlation
length convention d int r^(d-1) R dr = 1, detailed balance
sigma_PS = (d-1) K^d sigma_SP, the elastic channels' reduction to the
acoustic formula, and closed forms of Sigma for single-parameter media.
"""
from __future__ import annotations

import math
from typing import Callable, Dict, List

import numpy as np
from scipy import special

LAWS = ('exp', 'power_law', 'gaussian', 'triangular', 'low_pass', 'VonKarman')


def psdf(law: str, d: int, nu: float = 0.5):
    """(Phi, R, kinks) of a PSDF law in dimension d: the dimensionless spectrum,
    its correlation function, and the wavenumbers where Phi is not smooth
    (MaterialClass.m:479-710)."""
    if law == 'exp':                                                   # :493-508
        if d == 2:
            a = math.sqrt(2.0)
            return (lambda z: 1.0 / (4 * math.pi * (1 + (z / a) ** 2) ** 1.5),
                    lambda r: np.exp(-a * r), ())
        if d == 3:
            a = 6.0 ** (1 / 3)
            return (lambda z: 1.0 / (6 * math.pi ** 2 * (1 + (z / a) ** 2) ** 2),
                    lambda r: np.exp(-a * r), ())
    if law == 'power_law':                                             # :535-552
        if d == 2:
            def phi2(z):
                # z K1(z) / (4 pi) as written: 0 * Inf = NaN at z = 0 (the limit is
                # 1/(4 pi)); prepareSigmaOne's cumsum(...,'omitnan') then counts 0 there
                z = np.asarray(z, dtype=np.float64)
                with np.errstate(invalid='ignore'):
                    return z * special.k1(z) / (4 * math.pi)
            return phi2, (lambda r: 1.0 / (1 + r ** 2) ** 2), ()
        if d == 3:
            a = (3 * math.pi / 4) ** (1 / 3)
            return (lambda z: np.exp(-np.asarray(z) / a) / (6 * math.pi ** 2),
                    lambda r: 1.0 / (1 + (a * r) ** 2) ** 2, ())
    if law == 'gaussian':                                              # :575-589
        if d == 2:
            return (lambda z: np.exp(-np.asarray(z) ** 2 / 4) / (4 * math.pi),
                    lambda r: np.exp(-np.asarray(r) ** 2), ())
        if d == 3:
            a = (3 * math.sqrt(math.pi) / 4) ** (2 / 3)
            return (lambda z: np.exp(-np.asarray(z) ** 2 / (4 * a)) / (6 * math.pi ** 2),
                    lambda r: np.exp(-a * np.asarray(r) ** 2), ())
    if law == 'triangular' and d == 3:                                 # :620-623
        def rt(r):
            x = 2 * math.pi * np.asarray(r, dtype=np.float64)
            return 12 * (2 - 2 * np.cos(x) - x * np.sin(x)) / x ** 4
        return (lambda z: (3 / (8 * math.pi ** 4)) * (1 - np.asarray(z) / (2 * math.pi))
                * (np.asarray(z) < 2 * math.pi), rt, (2 * math.pi,))
    if law == 'low_pass' and d == 3:                                   # :652-655
        a = 3 * math.pi / 2

        def rl(r):
            x = a * np.asarray(r, dtype=np.float64)
            return 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
        return (lambda z: (2 / (9 * math.pi ** 4)) * (np.asarray(z) < a), rl, (a,))
    if law == 'VonKarman':                                             # :693-705
        if d == 2:
            a = 2 * math.sqrt(nu)
            c = 1 / (4 * math.pi)
            e = nu + 1
        elif d == 3:
            a = (6 * math.sqrt(math.pi) * math.gamma(nu + 1.5) / math.gamma(nu)) ** (1 / 3)
            c = 1 / (6 * math.pi ** 2)
            e = nu + 1.5
        else:
            raise ValueError(d)

        def rv(r, a=a):
            x = a * np.asarray(r, dtype=np.float64)
            return 2 ** (1 - nu) / math.gamma(nu) * x ** nu * special.kv(nu, x)
        return (lambda z, a=a, c=c, e=e: c / (1 + (np.asarray(z) / a) ** 2) ** e), rv, ()
    raise ValueError(f"no {law} law in {d}-D")


class Medium:
    """The scattering model of one isotropic random medium (the MaterialClass
    properties calcSigmaIsotropic reads)."""

    def __init__(self, d: int, acoustics: bool, freq: float, v, cv, corr, law: str, lc: float, nu: float = 0.5):
        self.d, self.acoustics, self.freq, self.law, self.lc, self.nu = d, bool(acoustics), float(freq), law, lc, nu
        self.v = np.atleast_1d(np.asarray(v, dtype=np.float64))
        self.cv = np.atleast_1d(np.asarray(cv, dtype=np.float64))
        self.corr = np.atleast_1d(np.asarray(corr, dtype=np.float64))
        self.Phi, self.R, self.kinks = psdf(law, d, nu)

    # -- differential cross-sections (MaterialClass.m:200-357) ------------
    def channels(self) -> List[List[Callable]]:
        """[[sigma]] (acoustic) or [[PP, PS], [SP, SS]] (elastic) with their
        PSDF arguments q(theta) = zeta * sqrt(A - B cos theta)."""
        d, om = self.d, 2 * math.pi * self.freq
        Phi = self.Phi
        if self.acoustics:                                             # :202-236
            zeta = om / self.v[0] * self.lc
            dk, dr = self.cv[0], self.cv[1]
            rho = self.corr[0]

            def s(th):
                c = np.cos(th)
                return (math.pi / 2) * om * zeta ** d * (c ** 2 * dr ** 2 + 2 * c * rho * dk * dr + dk ** 2) \
                    * Phi(zeta * np.sqrt(2 * (1 - c)))
            s.q = (zeta, 2.0, 2.0)
            return [[s]]
        vp, vs = self.v[0], self.v[1]
        K = vp / vs
        zp = om / vp * self.lc                                         # :244-248
        zs = K * zp
        dl, dm, dr = self.cv[0], self.cv[1], self.cv[2]
        rlm, rlr, rmr = self.corr[0], self.corr[1], self.corr[2]

        def pp(th):                                                    # :275-282, :328-333
            c = np.cos(th)
            return (math.pi / 2) * om * zp ** d * (
                (1 - 2 / K ** 2) ** 2 * dl ** 2 + 4 * (1 / K ** 2 - 2 / K ** 4) * rlm * dl * dm * c ** 2
                + (4 / K ** 4) * dm ** 2 * c ** 4 + dr ** 2 * c ** 2 - 2 * (1 - 2 / K ** 2) * rlr * dl * dr * c
                - (4 / K ** 2) * rmr * dm * dr * c ** 3) * Phi(zp * np.sqrt(2 * (1 - c)))
        pp.q = (zp, 2.0, 2.0)
        if d == 2:
            def ps(th):                                                # :285-290
                c = np.cos(th)
                return (math.pi / 2) * om * zp ** 2 * (K ** 2 * dr ** 2 + 4 * dm ** 2 * c ** 2
                                                       - 4 * K * rmr * dm * dr * c) \
                    * (1 - c ** 2) * Phi(zp * np.sqrt(1 + K ** 2 - 2 * K * c))

            def sp(th):                                                # :293-298
                c = np.cos(th)
                return (math.pi / 2) * om * zp ** 2 * (dr ** 2 + (4 / K ** 2) * dm ** 2 * c ** 2
                                                       - (4 / K) * rmr * dm * dr * c) \
                    * (1 - c ** 2) * Phi(zs * np.sqrt(1 + 1 / K ** 2 - 2 * c / K))

            def ss(th):                                                # :301-307
                c = np.cos(th)
                g = 2 * c ** 2 - 1
                return (math.pi / 2) * om * zs ** 2 * (dr ** 2 * c ** 2 + dm ** 2 * g ** 2
                                                       - 2 * rmr * dm * dr * c * g) \
                    * Phi(zs * np.sqrt(2 * (1 - c)))
        else:
            def ps(th):                                                # :336-338
                c = np.cos(th)
                return (math.pi / 2) * K * om * zp ** 3 * (K ** 2 * dr ** 2 + 4 * dm ** 2 * c ** 2
                                                           - 4 * K * rmr * dm * dr * c) \
                    * (1 - c ** 2) * Phi(zp * np.sqrt(1 + K ** 2 - 2 * K * c))

            def sp(th):                                                # :341-343
                c = np.cos(th)
                return (math.pi / 4 / K ** 3) * om * zs ** 3 * (dr ** 2 + (4 / K ** 2) * dm ** 2 * c ** 2
                                                                 - (4 / K) * rmr * dm * dr * c) \
                    * (1 - c ** 2) * Phi(zs * np.sqrt(1 + 1 / K ** 2 - 2 / K * c))

            def ss(th):                                                # :346-353
                c = np.cos(th)
                f = Phi(zs * np.sqrt(2 * (1 - c)))
                return (math.pi / 4) * om * zs ** 3 * (dr ** 2 * (1 + c ** 2) + dm ** 2 * (4 * c ** 4 - 3 * c ** 2 + 1)
                                                       - rmr * dm * dr * 4 * c ** 3) * f
        ps.q = (zp, 1 + K ** 2, 2 * K)
        sp.q = (zs, 1 + 1 / K ** 2, 2 / K)
        ss.q = (zs, 2.0, 2.0)
        return [[pp, ps], [sp, ss]]


# -- quadrature ---------------------------------------------------------

_GL_X, _GL_W = np.polynomial.legendre.leggauss(24)


def _breaks(s, kinks) -> np.ndarray:
    """Angles in (0, pi) where sigma's PSDF argument crosses a kink of Phi."""
    zeta, A, B = s.q
    out = [0.0, math.pi]
    for k in kinks:
        c = (A - (k / zeta) ** 2) / B
        if -1 < c < 1:
            out.append(math.acos(c))
    return np.unique(out)


def integrate(f: Callable, s, kinks, panels: int = 400) -> float:
    """int_0^pi f(theta) dtheta by composite 24-point Gauss-Legendre over
    `panels` panels per smooth piece."""
    br = _breaks(s, kinks)
    tot = 0.0
    for a, b in zip(br[:-1], br[1:]):
        e = np.linspace(a, b, panels + 1)
        h = np.diff(e) / 2
        mid = (e[:-1] + e[1:]) / 2
        x = mid[:, None] + h[:, None] * _GL_X[None, :]
        tot += float(np.sum(h[:, None] * _GL_W[None, :] * f(x)))
    return tot


def sigma_one(s, d: int, kinks=()) -> tuple:
    """(Sigma, Sigma') of one channel (MaterialClass.m:2446-2453)."""
    if d == 2:
        return 2 * integrate(s, s, kinks), 2 * integrate(lambda t: s(t) * np.cos(t), s, kinks)
    return (2 * math.pi * integrate(lambda t: s(t) * np.sin(t), s, kinks),
            2 * math.pi * integrate(lambda t: s(t) * np.sin(t) * np.cos(t), s, kinks))


def knots(s, d: int, Sigma: float, Nth: int = 1_000_000, threshold: float = 1e-8):
    """Inverse-CDF knots (cdf, theta) of prepareSigmaOne (MaterialClass.m:2443-2462)."""
    th = np.linspace(0.0, math.pi, Nth)
    if d == 2:
        pdf = (2 / Sigma) * s(th)
    else:
        pdf = (2 * math.pi / Sigma) * np.sin(th) * s(th)
    pdf = np.where(np.isnan(pdf), 0.0, pdf)                            # cumsum 'omitnan'
    cdf = np.cumsum(pdf) * (math.pi / Nth)
    ind = np.nonzero(np.diff(cdf) > threshold)[0]
    ind = np.unique(np.concatenate([ind, ind + 1]))
    return cdf[ind], th[ind]


def prepare(m: Medium, with_tables: bool = True) -> Dict:
    """prepareSigma (MaterialClass.m:2379-2441): Sigma, Sigma', mean free
    times, P2P / S2S and (optionally) the knots of every channel."""
    ch = m.channels()
    n = len(ch)
    S = np.zeros((n, n))
    Sp = np.zeros((n, n))
    tabs = [[None] * n for _ in range(n)]
    for a in range(n):
        for b in range(n):
            S[a, b], Sp[a, b] = sigma_one(ch[a][b], m.d, m.kinks)
            if with_tables:
                tabs[a][b] = knots(ch[a][b], m.d, S[a, b])
    out = {'Sigma': S, 'Sigmapr': Sp, 'tables': tabs}
    if m.acoustics:
        out['mft'] = np.array([1 / S[0, 0]])
        out['P2P'] = out['S2S'] = 1.0
    else:
        out['mft'] = 1 / S.sum(axis=1)                                 # :2421
        out['P2P'] = S[0, 0] / S[0].sum()                              # :2439
        out['S2S'] = S[1, 1] / S[1].sum()                              # :2440
    return out

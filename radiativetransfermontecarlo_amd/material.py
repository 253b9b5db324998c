"""MaterialClass: scattering cross-sections and the tables the kernels read.

Host-side restatement of the parts of ``@MaterialClass/MaterialClass.m`` that
``radiativeTransfer`` uses: the constructor (MaterialClass.m:59-97), the
differential scattering cross-sections of ``calcSigmaIsotropic``
(:172-359), the PSDF laws (:479-710), ``prepareSigma`` / ``prepareSigmaOne``
(:2379-2463), ``preset`` (:1765-1785) and ``Zoeppritz`` with the solid/air
``ZoeppritzFluid`` (MaterialClass.m:1786-1885, ZoeppritzFluid.m:1-245).

Property names follow the MATLAB class so code written against it reads the
same.  ``sigma`` entries are Python callables of a numpy angle array.
"""
from __future__ import annotations

import copy
import math
import warnings
from typing import Callable, List, Optional

import numpy as np
from scipy import integrate, special

__all__ = ["MaterialClass", "InverseCDF", "zoeppritz_fluid"]

_SPECTRAL_LAWS = ['', 'exp', 'power_law', 'gaussian', 'triangular', 'low_pass', 'VonKarman',
                  'monodispersesphere', 'image', 'Imported']


def _validatestring(value: str, options: List[str]) -> str:
    """MATLAB validatestring: case-insensitive, unambiguous prefix match (MaterialClass.m:138-140)."""
    v = str(value)
    exact = [o for o in options if o.lower() == v.lower()]
    if exact:
        return exact[0]
    cands = [o for o in options if o and o.lower().startswith(v.lower())]
    if len(cands) == 1 and v:
        return cands[0]
    raise ValueError(f"Expected input to match one of these values: {', '.join(repr(o) for o in options)}; "
                     f"the input, '{value}', did not match any valid string (or matched several).")


class InverseCDF:
    """``griddedInterpolant(x, y)`` with linear interpolation and linear (default)
    or nearest extrapolation: the object ``prepareSigmaOne`` stores in ``invcdf``.

    ``x`` are the CDF knots (``GridVectors{1}``), ``y`` the angles (``Values``).
    """

    def __init__(self, x: np.ndarray, y: np.ndarray, extrap: str = "linear"):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        if x.ndim != 1 or x.shape != y.shape or x.size < 2:
            raise ValueError("griddedInterpolant needs matching vectors of >= 2 knots")
        if not np.all(np.diff(x) > 0):
            raise ValueError("The grid vectors must be strictly monotonic increasing.")
        self.x = x
        self.y = y
        self.extrap = extrap

    @property
    def GridVectors(self):
        return (self.x,)

    @property
    def Values(self):
        return self.y

    def __call__(self, q):
        q = np.asarray(q, dtype=np.float64)
        x, y, n = self.x, self.y, self.x.size
        k = np.searchsorted(x, q, side="right") - 1
        k = np.clip(k, 0, n - 2)
        t = (q - x[k]) / (x[k + 1] - x[k])
        out = y[k] + t * (y[k + 1] - y[k])
        if self.extrap == "nearest":
            out = np.where(q < x[0], y[0], out)
            out = np.where(q > x[-1], y[-1], out)
        return out


def _heaviside(h):
    """MaterialClass.heaviside (MaterialClass.m:1760-1764): H(0) = 1/2."""
    h = np.asarray(h, dtype=np.float64)
    return np.where(h > 0, 1.0, np.where(h == 0, 0.5, 0.0))


def _integral(f: Callable, a: float, b: float) -> float:
    """MATLAB ``integral`` (adaptive Gauss-Kronrod); QUADPACK at tighter tolerance."""
    val, _err = integrate.quad(lambda t: float(np.asarray(f(np.array([t])))[0]), a, b,
                               epsabs=1e-13, epsrel=1e-12, limit=500)
    return val


class MaterialClass:
    """Material description and scattering operator (MaterialClass.m:10-97)."""

    def __init__(self, geometry=None, freq=None, acoustics=None, v=None,
                 coefficients_of_variation=None, correlation_coefficients=None, acf=None, lc=None):
        self.d = 3
        self.type = 'isotropic'
        self.acoustics = None
        self.v = None
        self.vp = None
        self.vs = None
        self.rho = None
        self.Frequency = None
        self.Q = math.inf
        self.coefficients_of_variation = None
        self.correlation_coefficients = None
        self._SpectralLaw = ''
        self.SpectralParam = {}
        self.CorrelationLength = None
        self.sigma: list = []
        self.Sigma = None
        self.Sigmapr = None
        self.invcdf = None
        self.Diffusivity = None
        self.meanFreeTime = None
        self.meanFreePath = None
        self.transportMeanFreeTime = None
        self.transportMeanFreePath = None
        self.g = None
        self.P2P = None
        self.S2S = None
        self.Phi = None
        self.R = None
        self.timeSteps = 0
        if geometry is not None:
            dim = geometry["dimension"] if isinstance(geometry, dict) else geometry.dimension
            if dim == 1:
                warnings.warn("Total scattering cross-sections for 1D case are to be implemented in "
                              "future releases of the code")
            self.d = int(dim)
            self.acoustics = bool(acoustics)
            self.Frequency = freq
            if acoustics:
                self.v = float(np.asarray(v).ravel()[0])
            else:
                vv = np.asarray(v, dtype=np.float64).ravel()
                self.vp = float(vv[0])
                self.vs = float(vv[1])
            self.coefficients_of_variation = np.asarray(coefficients_of_variation, dtype=np.float64).ravel()
            self.correlation_coefficients = np.asarray(correlation_coefficients, dtype=np.float64).ravel()
            self.CorrelationLength = lc
            self.SpectralLaw = acf
            self.timeSteps = 0

    # -- properties -----------------------------------------------------
    @property
    def SpectralLaw(self):
        return self._SpectralLaw

    @SpectralLaw.setter
    def SpectralLaw(self, value):
        self._SpectralLaw = _validatestring(value, _SPECTRAL_LAWS)

    def copyobj(self):
        return copy.deepcopy(self)

    # -- scattering cross-sections (MaterialClass.m:141-359) ------------
    def CalcSigma(self):
        if self.Phi is None:
            self.getPSDF()
        if self.type == 'isotropic':
            self.calcSigmaIsotropic()
        else:
            raise NotImplementedError('DSCS type not supported')

    def calcSigmaIsotropic(self):
        Phi = self.Phi
        if self.acoustics:
            omega = 2 * np.pi * self.Frequency
            zeta = omega / self.v * self.CorrelationLength                       # :204
            dkk = self.coefficients_of_variation[0]
            drr = self.coefficients_of_variation[1]
            rkr = float(np.asarray(self.correlation_coefficients).ravel()[0])
            if self.d == 1:
                raise NotImplementedError('sigma for 1D not implemented')
            p = 2 if self.d == 2 else 3

            def s(th, omega=omega, zeta=zeta, dkk=dkk, drr=drr, rkr=rkr, p=p):     # :225-235
                th = np.asarray(th, dtype=np.float64)
                c = np.cos(th)
                return (np.pi / 2) * omega * zeta ** p * (c ** 2 * drr ** 2 + 2 * c * rkr * dkk * drr + dkk ** 2) \
                    * Phi(zeta * np.sqrt(2 * (1 - c)))
            self.sigma = [s]
            return
        K = self.vp / self.vs                                                     # :242-253
        omega = 2 * np.pi * self.Frequency
        zP = omega / self.vp * self.CorrelationLength
        zS = K * zP
        dll, dmm, drr = (float(a) for a in self.coefficients_of_variation[:3])
        rlm, rlr, rmr = (float(a) for a in self.correlation_coefficients[:3])
        if self.d == 1:
            raise NotImplementedError('sigma for 1D not implemented')
        if self.d == 2:                                                            # :259-310
            def sPP(th):
                c = np.cos(np.asarray(th, dtype=np.float64))
                return (np.pi / 2) * omega * zP ** 2 * (
                    (1 - 2 / K ** 2) ** 2 * dll ** 2 + 4 * (1 / K ** 2 - 2 / K ** 4) * rlm * dll * dmm * c ** 2
                    + (4 / K ** 4) * dmm ** 2 * c ** 4 + drr ** 2 * c ** 2
                    - 2 * (1 - 2 / K ** 2) * rlr * dll * drr * c - (4 / K ** 2) * rmr * dmm * drr * c ** 3) \
                    * Phi(zP * np.sqrt(2 * (1 - c)))

            def sPS(th):
                c = np.cos(np.asarray(th, dtype=np.float64))
                return (np.pi / 2) * omega * zP ** 2 * (K ** 2 * drr ** 2 + 4 * dmm ** 2 * c ** 2
                                                        - 4 * K * rmr * dmm * drr * c) \
                    * (1 - c ** 2) * Phi(zP * np.sqrt(1 + K ** 2 - 2 * K * c))

            def sSP(th):
                c = np.cos(np.asarray(th, dtype=np.float64))
                return (np.pi / 2) * omega * zP ** 2 * (drr ** 2 + (4 / K ** 2) * dmm ** 2 * c ** 2
                                                        - (4 / K) * rmr * dmm * drr * c) \
                    * (1 - c ** 2) * Phi(zS * np.sqrt(1 + 1 / K ** 2 - 2 * c / K))

            def sSS(th):
                c = np.cos(np.asarray(th, dtype=np.float64))
                G = 2 * c ** 2 - 1
                return (np.pi / 2) * omega * zS ** 2 * (drr ** 2 * c ** 2 + dmm ** 2 * G ** 2
                                                        - 2 * rmr * dmm * drr * c * G) \
                    * Phi(zS * np.sqrt(2 * (1 - c)))
            self.sigma = [[sPP, sPS], [sSP, sSS]]
            return

        def sPP3(th):                                                              # :328-333
            c = np.cos(np.asarray(th, dtype=np.float64))
            return (np.pi / 2) * omega * zP ** 3 * (
                (1 - 2 / K ** 2) ** 2 * dll ** 2 + 4 * (1 / K ** 2 - 2 / K ** 4) * rlm * dll * dmm * c ** 2
                + (4 / K ** 4) * dmm ** 2 * c ** 4 + drr ** 2 * c ** 2
                - 2 * (1 - 2 / K ** 2) * rlr * dll * drr * c - (4 / K ** 2) * rmr * dmm * drr * c ** 3) \
                * Phi(zP * np.sqrt(2 * (1 - c)))

        def sPS3(th):                                                              # :336-338
            c = np.cos(np.asarray(th, dtype=np.float64))
            return (np.pi / 2) * K * omega * zP ** 3 * (K ** 2 * drr ** 2 + 4 * dmm ** 2 * c ** 2
                                                        - 4 * K * rmr * dmm * drr * c) \
                * (1 - c ** 2) * Phi(zP * np.sqrt(1 + K ** 2 - 2 * K * c))

        def sSP3(th):                                                              # :341-343
            c = np.cos(np.asarray(th, dtype=np.float64))
            return (np.pi / 4 / K ** 3) * omega * zS ** 3 * (drr ** 2 + (4 / K ** 2) * dmm ** 2 * c ** 2
                                                             - (4 / K) * rmr * dmm * drr * c) \
                * (1 - c ** 2) * Phi(zS * np.sqrt(1 + 1 / K ** 2 - 2 / K * c))

        def sSS3(th):                                                              # :346-353
            c = np.cos(np.asarray(th, dtype=np.float64))
            ph = Phi(zS * np.sqrt(2 * (1 - c)))
            tt = (np.pi / 4) * omega * zS ** 3 * drr ** 2 * (1 + c ** 2) * ph
            gg = (np.pi / 4) * omega * zS ** 3 * dmm ** 2 * (4 * c ** 4 - 3 * c ** 2 + 1) * ph
            gt = -(np.pi / 4) * omega * zS ** 3 * rmr * dmm * drr * (4 * c ** 3) * ph
            return tt + gg + gt
        self.sigma = [[sPP3, sPS3], [sSP3, sSS3]]

    # -- power spectral densities (MaterialClass.m:381-710) --------------
    def getPSDF(self):
        if self.acoustics:
            rkr = float(np.asarray(self.correlation_coefficients).ravel()[0])
            if abs(rkr) > 1:
                raise ValueError('Absolute value of corr(kappa,rho) should be less than 1!')
        else:
            rlm, rlr, rmr = (float(a) for a in np.asarray(self.correlation_coefficients).ravel()[:3])
            C = np.array([[1, rlm, rlr], [rlm, 1, rmr], [rlr, rmr, 1]])
            if np.min(np.linalg.eigvalsh(C)) < -1e-12:
                raise ValueError('The lambda-mu-rho correlation matrix is not positive semidefinite.')
            if np.any(np.abs(self.correlation_coefficients) > 1):
                raise ValueError('Absolute values of correlation coefficients should be less than 1!')
        law = self.SpectralLaw
        lc = self.CorrelationLength
        if law == 'exp':
            self.Exponential(lc)
        elif law == 'power_law':
            self.PowerLaw(lc)
        elif law == 'gaussian':
            self.Gaussian(lc)
        elif law == 'triangular':
            self.Triangular(lc)
        elif law == 'low_pass':
            self.LowPass(lc)
        elif law == 'VonKarman':
            if 'nu' not in (self.SpectralParam or {}):
                raise ValueError('Please add the Hurst number for VonKarman PSDF')
            self.VonKarman(lc, self.SpectralParam['nu'])
        elif law in ('monodispersesphere', 'image', 'Imported'):
            raise NotImplementedError(f"SpectralLaw '{law}' (microstructure PSDF) is not part of this build")

    def Exponential(self, lc):
        self.SpectralParam = {'correlationLength': lc}
        self.CorrelationLength = lc
        self.SpectralLaw = 'Exp'
        if self.d == 1:
            self.R = lambda z: np.exp(-2 * z)
            self.Phi = lambda z: 1.0 / (2 * np.pi * (1 + (np.asarray(z) / 2) ** 2))
        elif self.d == 2:
            a = math.sqrt(2)
            self.R = lambda z: np.exp(-a * z)
            self.Phi = lambda z: 1.0 / (4 * np.pi * (1 + (np.asarray(z) / a) ** 2) ** 1.5)
        elif self.d == 3:
            a = 6 ** (1 / 3)
            self.R = lambda z: np.exp(-a * z)
            self.Phi = lambda z: 1.0 / (6 * np.pi ** 2 * (1 + (np.asarray(z) / a) ** 2) ** 2)
        else:
            raise ValueError('incorrect dimension!')
        return self.Phi

    def PowerLaw(self, lc):
        self.SpectralParam = {'correlationLength': lc}
        self.CorrelationLength = lc
        self.SpectralLaw = 'power_law'
        if self.d == 1:
            a = np.pi / 2
            self.R = lambda z: 1.0 / (1 + (a * z) ** 2) ** 2
            self.Phi = lambda z: 1.0 / (2 * np.pi) * (1 + np.asarray(z) / a) * np.exp(-np.asarray(z) / a)
        elif self.d == 2:
            self.R = lambda z: 1.0 / (1 + z ** 2) ** 2
            self.Phi = lambda z: np.asarray(z) * special.kv(1, z) / (4 * np.pi)
        elif self.d == 3:
            a = (3 * np.pi / 4) ** (1 / 3)
            self.R = lambda z: 1.0 / (1 + (a * z) ** 2) ** 2
            self.Phi = lambda z: 1.0 / (6 * np.pi ** 2) * np.exp(-np.asarray(z) / a)
        return self.Phi

    def Gaussian(self, lc):
        self.SpectralParam = {'correlationLength': lc}
        self.CorrelationLength = lc
        self.SpectralLaw = 'Gauss'
        if self.d == 1:
            self.R = lambda z: np.exp(-np.pi * z ** 2)
            self.Phi = lambda z: 1.0 / (2 * np.pi) * np.exp(-np.asarray(z) ** 2 / (4 * np.pi))
        elif self.d == 2:
            self.R = lambda z: np.exp(-z ** 2)
            self.Phi = lambda z: 1.0 / (4 * np.pi) * np.exp(-np.asarray(z) ** 2 / 4)
        elif self.d == 3:
            a = (3 * math.sqrt(np.pi) / 4) ** (2 / 3)
            self.R = lambda z: np.exp(-a * z ** 2)
            self.Phi = lambda z: 1.0 / (6 * np.pi ** 2) * np.exp(-np.asarray(z) ** 2 / (4 * a))
        else:
            raise ValueError('incorrect dimension!')
        return self.Phi

    def Triangular(self, lc):
        self.SpectralParam = {'correlationLength': lc}
        self.CorrelationLength = lc
        self.SpectralLaw = 'Triangular'
        if self.d == 3:
            self.R = lambda z: (12 * (2 - 2 * np.cos(2 * np.pi * z) - (2 * np.pi * z) * np.sin(2 * np.pi * z))) \
                / (2 * np.pi * z) ** 4
            self.Phi = lambda z: (3 / 8 / np.pi ** 4) * (1 - np.asarray(z) / 2 / np.pi) \
                * _heaviside(2 * np.pi - np.asarray(z))
        else:
            raise ValueError('dimensions other than 3 are not coded yet!')
        return self.Phi

    def LowPass(self, lc):
        self.SpectralParam = {'correlationLength': lc}
        self.CorrelationLength = lc
        self.SpectralLaw = 'low_pass'
        if self.d == 3:
            a = 3 * np.pi / 2
            self.R = lambda z: (3 * (np.sin(a * z) - a * z * np.cos(a * z))) / (a * z) ** 3
            self.Phi = lambda z: (2 / 9 / np.pi ** 4) * _heaviside(a - np.asarray(z))
        else:
            raise ValueError('LowPass spectral model is currently implemented only for d = 3.')
        return self.Phi

    def VonKarman(self, lc, nu):
        self.SpectralParam = {'correlationLength': lc, 'nu': nu}
        self.SpectralLaw = 'VonKarman'
        self.CorrelationLength = lc
        g = math.gamma
        if self.d == 1:
            a = 2 * math.sqrt(np.pi) * g(nu + 0.5) / g(nu)
            self.Phi = lambda z: 1 / (2 * np.pi) / (1 + (np.asarray(z) / a) ** 2) ** (nu + 0.5)
        elif self.d == 2:
            a = 2 * math.sqrt(nu)
            self.Phi = lambda z: 1 / (4 * np.pi) / (1 + (np.asarray(z) / a) ** 2) ** (nu + 1)
        elif self.d == 3:
            a = (6 * math.sqrt(np.pi) * g(nu + 1.5) / g(nu)) ** (1 / 3)
            self.Phi = lambda z: 1 / (6 * np.pi ** 2) / (1 + (np.asarray(z) / a) ** 2) ** (nu + 1.5)
        else:
            raise ValueError('incorrect dimension!')
        self.R = lambda z: 2 ** (1 - nu) / g(nu) * (a * z) ** nu * special.kv(nu, a * z)
        return self.Phi

    # -- presets (MaterialClass.m:1765-1785) -----------------------------
    @staticmethod
    def preset(n: int) -> "MaterialClass":
        obj = MaterialClass()
        if n == 1:
            obj.sigma = [lambda th: 1 / 4 / np.pi * np.ones(np.shape(th))]
            obj.v = 2
            obj.acoustics = True
        elif n == 2:
            obj.sigma = [lambda th: 1 / 10 / np.pi * np.ones(np.shape(th))]
            obj.v = 2
            obj.acoustics = True
        elif n in (3, 4):
            obj.vp = 6
            obj.vs = 6 / math.sqrt(3)
            obj.acoustics = False
        return obj

    # -- scattering tables (MaterialClass.m:2379-2463) -------------------
    @staticmethod
    def prepareSigmaOne(sigma: Callable, d: int, Nth: int = 1_000_000, knot_threshold: float = 1e-8):
        """Sigma, Sigma', and the reference's inverse CDF of the scattering angle.

        Knots are kept where the cdf step exceeds ``knot_threshold`` (the
        reference's 1e-8, MaterialClass.m:2460) and the table extrapolates
        linearly beyond them, so the sampled tail beyond the last knot is not
        sigma's.  ``knot_threshold=0`` keeps every strictly increasing knot:
        the table then follows sigma to within the 1e6-point grid everywhere,
        which is what the adjoint's likelihood-ratio score assumes (beyond
        the last knot it then holds the last angle)."""
        xth = np.linspace(0, np.pi, Nth)
        d = int(d)
        if d == 2:
            Sigma = 2 * _integral(sigma, 0, np.pi)
            Sigma_prime = 2 * _integral(lambda th: sigma(th) * np.cos(th), 0, np.pi)
            pdf = (2 / Sigma) * np.asarray(sigma(xth), dtype=np.float64)
        elif d == 3:
            Sigma = 2 * np.pi * _integral(lambda th: sigma(th) * np.sin(th), 0, np.pi)
            Sigma_prime = 2 * np.pi * _integral(lambda th: sigma(th) * np.sin(th) * np.cos(th), 0, np.pi)
            pdf = (2 * np.pi / Sigma) * np.sin(xth) * np.asarray(sigma(xth), dtype=np.float64)
        else:
            raise ValueError("prepareSigmaOne: d must be 2 or 3")
        pdf = np.broadcast_to(pdf, xth.shape).astype(np.float64)
        if np.any(np.isnan(pdf)):
            warnings.warn('There is a NaN inside the probability density function')
        cdf = np.nancumsum(pdf) * (np.pi / Nth)
        step = np.diff(cdf) > knot_threshold
        keep = np.zeros(Nth, dtype=bool)          # i where step i or step i-1 is kept
        keep[:-1] |= step
        keep[1:] |= step
        ind = np.nonzero(keep)[0]
        if knot_threshold <= 0:
            # cdf steps that underflow next to 1 would repeat a knot; beyond the
            # last knot (mass ~1/Nth) the table holds its last angle instead of
            # extrapolating the near-vertical last segment
            ind = ind[np.concatenate([[True], np.diff(cdf[ind]) > 0])]
            return Sigma, Sigma_prime, InverseCDF(cdf[ind], xth[ind], extrap="nearest")
        invcdf = InverseCDF(cdf[ind], xth[ind])
        return Sigma, Sigma_prime, invcdf

    @staticmethod
    def prepareSigma(mat: "MaterialClass", d: int, knot_threshold: float = 1e-8) -> "MaterialClass":
        if not mat.sigma:
            mat.CalcSigma()
        d = int(d)
        if mat.acoustics:
            mat.Sigma, mat.Sigmapr, mat.invcdf = MaterialClass.prepareSigmaOne(mat.sigma[0], d,
                                                                                    knot_threshold=knot_threshold)
            mat.Diffusivity = mat.v ** 2 / (float(d) * (mat.Sigma - mat.Sigmapr))
            mat.meanFreeTime = 1 / mat.Sigma
            mat.meanFreePath = mat.v * mat.meanFreeTime
            mat.transportMeanFreePath = float(d) * mat.Diffusivity / mat.v
            mat.transportMeanFreeTime = mat.transportMeanFreePath / mat.v
            mat.g = 1 - mat.meanFreePath / mat.transportMeanFreePath
            mat.vp = mat.v
            mat.vs = 0
            return mat
        S = np.zeros((2, 2))
        Sp = np.zeros((2, 2))
        inv = [[None, None], [None, None]]
        for a in range(2):
            for b in range(2):
                S[a, b], Sp[a, b], inv[a][b] = MaterialClass.prepareSigmaOne(mat.sigma[a][b], d,
                                                                                 knot_threshold=knot_threshold)
        mat.Sigma, mat.Sigmapr, mat.invcdf = S, Sp, inv
        mat.meanFreeTime = 1.0 / S.sum(axis=1)
        vpvs = np.array([mat.vp, mat.vs], dtype=np.float64)
        mat.meanFreePath = vpvs * mat.meanFreeTime
        K = mat.vp / mat.vs
        den = (S[0, 0] + S[0, 1] - Sp[0, 0]) * (S[1, 1] + S[1, 0] - Sp[1, 1]) - Sp[0, 1] * Sp[1, 0]
        tP = (mat.vp * (S[1, 1] + S[1, 0] - Sp[1, 1]) + mat.vs * Sp[0, 1]) / den
        tS = (mat.vs * (S[0, 0] + S[0, 1] - Sp[0, 0]) + mat.vp * Sp[1, 0]) / den
        mat.transportMeanFreePath = np.array([tP, tS])
        # MATLAB's column / column is a least-squares division (MaterialClass.m:2431)
        mat.transportMeanFreeTime = float(np.dot(mat.transportMeanFreePath, vpvs) / np.dot(vpvs, vpvs))
        Dp = mat.vp * tP / d
        Ds = mat.vs * tS / d
        mat.Diffusivity = float((Dp + 2 * K ** 3 * Ds) / (1 + 2 * K ** 3))
        mat.P2P = S[0, 0] / S[0, :].sum()
        mat.S2S = S[1, 1] / S[1, :].sum()
        return mat

    # -- elastic interface (MaterialClass.m:1786-1846) --------------------
    @staticmethod
    def Zoeppritz(mat: "MaterialClass"):
        """Solid/air reflection coefficients on 181 angles 0..90 degrees."""
        if mat.acoustics:
            raise ValueError("This dont work for acoustic material")
        if mat.rho is None:
            raise ValueError('Please define the density in MaterialClass before calling MaterialInterface.')
        j1 = np.linspace(0, 90, 181)
        vp1, vs1, rho1 = float(mat.vp), float(mat.vs), float(mat.rho)
        vp2, vs2, rho2 = 343.0, 0.0, 1.0
        if vs2 != 0:
            raise NotImplementedError("ZoeppritzSolid is not reached: the interface fluid is air")
        out = zoeppritz_fluid(j1, vp1, vs1, rho1, vp2, vs2, rho2)
        out['j1_deg'] = j1
        return out


def _casin(x):
    return np.lib.scimath.arcsin(x)


def zoeppritz_fluid(theta, vp1, vs1, rho1, vp2, vs2, rho2) -> dict:
    """Amplitude and energy coefficients of a solid/fluid interface for P, SV
    and SH incidence (ZoeppritzFluid.m:1-245): 3x3 complex solves per angle."""
    theta = np.asarray(theta, dtype=np.float64)
    tr = theta * np.pi / 180
    n = theta.size
    keys = ['Rpp', 'Rpsv', 'Tpp', 'Tpsv', 'Rsp', 'Rsvsv', 'Tsp', 'Tsvsv', 'Rsh', 'Tsh']
    A = {k: np.zeros(n, dtype=np.complex128) for k in keys}
    Ek = ['E_Rpp', 'E_Rpsv', 'E_Tpp', 'E_Tpsv', 'E_Rsp', 'E_Rsvsv', 'E_Tsp', 'E_Tsvsv', 'E_Rsh', 'E_Tsh']
    E = {k: np.zeros(n) for k in Ek}

    def mat3(phi1, psi1, phi2, psi2):
        c1, cs1, c2 = np.cos(phi1), np.cos(psi1), np.cos(phi2)
        M = np.zeros((3, 3), dtype=np.complex128)
        M[0, 0] = c1
        M[0, 1] = -np.sin(psi1)
        M[0, 2] = c2
        M[1, 0] = -np.cos(2 * psi1)
        M[1, 1] = vs1 / vp1 * np.sin(2 * psi1)
        M[1, 2] = (rho2 * vp2 / (rho1 * vp1)) * np.cos(2 * psi2)
        M[2, 0] = np.sin(2 * phi1)
        M[2, 1] = vp1 / vs1 * np.cos(2 * psi1)
        M[2, 2] = (rho2 * vp1 * vs2 * vs2 / (rho1 * vp2 * vs1 * vs1)) * np.sin(2 * phi2)
        return M, c1, cs1, c2

    def solve(M, F):
        try:
            return np.linalg.solve(M, F)
        except np.linalg.LinAlgError:
            return np.full(3, np.nan + 0j)

    for i in range(n):
        # P incidence (:49-121)
        phi0 = tr[i]
        p = np.sin(phi0) / vp1
        phi1 = phi0 + 0j
        psi1 = _casin(vs1 * p + 0j)
        phi2 = _casin(vp2 * p + 0j)
        psi2 = _casin(vs2 * p + 0j)
        M, c1, cs1, c2 = mat3(phi1, psi1, phi2, psi2)
        psi0 = _casin(vs1 * np.sin(phi0) / vp1 + 0j)
        F = np.array([np.cos(phi0), np.cos(2 * psi0), np.sin(2 * phi0)], dtype=np.complex128)
        a = solve(M, F)
        A['Rpp'][i], A['Rpsv'][i], A['Tpp'][i] = a
        E['E_Rpp'][i] = abs(a[0]) ** 2
        E['E_Rpsv'][i] = (rho1 * vs1 * np.real(cs1)) / (rho1 * vp1 * np.real(c1)) * abs(a[1]) ** 2
        E['E_Tpp'][i] = (rho2 * vp2 * np.real(c2)) / (rho1 * vp1 * np.real(c1)) * abs(a[2]) ** 2
        if tr[i] == 0:
            E['E_Rpsv'][i] = 0.0
        # SV incidence (:123-188)
        psi0 = tr[i]
        p = np.sin(psi0) / vs1
        phi1 = _casin(vp1 * p + 0j)
        psi1 = psi0 + 0j
        phi2 = _casin(vp2 * p + 0j)
        psi2 = _casin(vs2 * p + 0j)
        M, c1, cs1, c2 = mat3(phi1, psi1, phi2, psi2)
        F = np.array([-np.sin(psi0), -vs1 * np.sin(2 * psi0) / vp1, vp1 * np.cos(2 * psi0) / vs1],
                     dtype=np.complex128)
        a = solve(M, F)
        A['Rsp'][i], A['Rsvsv'][i], A['Tsp'][i] = a
        E['E_Rsp'][i] = (rho1 * vp1 * np.real(c1)) / (rho1 * vs1 * np.real(cs1)) * abs(a[0]) ** 2
        E['E_Rsvsv'][i] = abs(a[1]) ** 2
        E['E_Tsp'][i] = (rho2 * vp2 * np.real(c2)) / (rho1 * vs1 * np.real(cs1)) * abs(a[2]) ** 2
        if tr[i] == 0:
            E['E_Rsp'][i] = 0.0
        # SH incidence (:190-221): total reflection at a free (fluid) interface
        A['Rsh'][i] = 1.0
        A['Tsh'][i] = 0.0
        E['E_Rsh'][i] = 1.0
        E['E_Tsh'][i] = 0.0
    out = dict(E)
    for k in keys:
        out['A_' + k] = A[k] if np.any(np.imag(A[k]) != 0) else np.real(A[k])
    return out

"""Native material preparation (rtmc_prepare_material): MaterialClass.prepareSigma
computed in librtmc (C++ on the host), for isotropic random media.

``params_from(material, d)`` maps a MaterialClass to ``rtmc_material_params``;
``prepare_host`` returns Sigma, Sigma', the mean free times, P2P / S2S and the
inverse-CDF tables as numpy arrays; ``NativeMaterial`` uploads them (and, for an
elastic material with a density, the solid/air Zoeppritz tables) for the
kernels (an ``rtmc_material`` usable wherever ``DeviceMaterial.c`` is).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from .material import MaterialClass

__all__ = ["params_from", "params_for", "prepare_host", "NativeMaterial", "NativeMaterials", "lc_derivatives",
           "zoeppritz_air", "compress_table", "KNOT_DTYPE"]

# rtmc_lds_knot as a numpy record
KNOT_DTYPE = np.dtype([("u", np.float64), ("theta", np.float32), ("slope", np.float32)])


def params_from(material: MaterialClass, d: int, knot_threshold: float = 1e-8, n_theta: int = 0,
                lds_tol: float = 0.0, lds_max_knots: int = 0):
    """rtmc_material_params of an (unprepared or prepared) isotropic MaterialClass.
    lds_tol > 0 also asks for compressed copies of the angle tables for the
    tracing kernel's LDS (rtmc_compress_table)."""
    if material.type != 'isotropic':
        raise ValueError("native preparation covers isotropic media (calcSigmaIsotropic)")
    law = material.SpectralLaw
    key = {'exp': 'exp', 'power_law': 'power_law', 'gaussian': 'gaussian', 'triangular': 'triangular',
           'low_pass': 'low_pass', 'VonKarman': 'VonKarman'}.get(law)
    if key is None:
        # prepared materials carry the law's display name (MaterialClass.m:479-710)
        key = {'Exp': 'exp', 'Gauss': 'gaussian', 'Triangular': 'triangular'}.get(law)
    if key is None:
        raise ValueError(f"SpectralLaw {law!r} has no native preparation")
    p = _abi.MaterialParams()
    p.dimension = int(d)
    p.acoustics = 1 if material.acoustics else 0
    if material.acoustics:
        p.v[0] = float(np.asarray(material.v).ravel()[0])
    else:
        p.v[0] = float(material.vp)
        p.v[1] = float(material.vs)
    p.frequency = float(np.asarray(material.Frequency).ravel()[0])
    cv = np.asarray(material.coefficients_of_variation, dtype=np.float64).ravel()
    cc = np.asarray(material.correlation_coefficients, dtype=np.float64).ravel()
    for k in range(min(3, cv.size)):
        p.cv[k] = cv[k]
    for k in range(min(3, cc.size)):
        p.corr[k] = cc[k]
    p.spectral_law = _abi.PSDF[key]
    p.lc = float(material.CorrelationLength)
    p.nu = float((material.SpectralParam or {}).get('nu', 0.0))
    Q = np.asarray(material.Q if material.Q is not None else math.inf, dtype=np.float64).ravel()
    p.Q[0] = Q[0]
    p.Q[1] = Q[1] if Q.size > 1 else Q[0]
    p.knot_threshold = float(knot_threshold)
    p.n_theta = int(n_theta)
    # an elastic density gives the solid/air tables of reflective boundaries
    rho = getattr(material, 'rho', None)
    p.rho = float(rho) if (rho is not None and not material.acoustics) else 0.0
    p.lds_tol = float(lds_tol)
    p.lds_max_knots = int(lds_max_knots)
    return p


def compress_table(x, y, tol: float = 1e-5, max_knots: int = 1024):
    """rtmc_compress_table on host arrays: (knots as a KNOT_DTYPE record array,
    the tolerance achieved)."""
    lib = _abi.load()
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    out = np.zeros(int(max_knots), dtype=KNOT_DTYPE)
    m = C.c_int32()
    used = C.c_double()
    P = C.POINTER(C.c_double)
    _abi.check(lib.rtmc_compress_table(x.ctypes.data_as(P), y.ctypes.data_as(P), x.size, float(tol), int(max_knots),
                                       out.ctypes.data_as(C.POINTER(_abi.LdsKnot)), C.byref(m), C.byref(used)),
               "rtmc_compress_table")
    return out[:m.value].copy(), used.value


def zoeppritz_air(vp: float, vs: float, rho: float, n: int = 181) -> dict:
    """rtmc_zoeppritz_air: E_Rpp and E_Rsvsv of a solid under air at
    linspace(0, 90, n) degrees (MaterialClass.Zoeppritz)."""
    lib = _abi.load()
    deg, rpp, rss = (np.zeros(n) for _ in range(3))
    P = C.POINTER(C.c_double)
    _abi.check(lib.rtmc_zoeppritz_air(float(vp), float(vs), float(rho), int(n), deg.ctypes.data_as(P),
                                      rpp.ctypes.data_as(P), rss.ctypes.data_as(P)), "rtmc_zoeppritz_air")
    return {'j1_deg': deg, 'E_Rpp': rpp, 'E_Rsvsv': rss}


def params_for(base: MaterialClass, d: int, c: float, lc: float, Q: float, knot_threshold: float = 1e-8,
               n_theta: int = 0, lds_tol: float = 0.0, lds_max_knots: int = 0):
    """params_from of ``base`` with the coefficients of variation scaled to
    cv[0] = c, the correlation length lc and the attenuation Q (acoustic)."""
    p = params_from(base, d, knot_threshold=knot_threshold, n_theta=n_theta, lds_tol=lds_tol,
                    lds_max_knots=lds_max_knots)
    s = float(c) / p.cv[0]
    for k in range(3):
        p.cv[k] *= s
    p.lc = float(lc)
    p.Q[0] = p.Q[1] = float(Q)
    return p


def prepare_host(params) -> dict:
    """rtmc_prepare_material_host as numpy arrays."""
    lib = _abi.load()
    h = _abi.MaterialHost()
    _abi.check(lib.rtmc_prepare_material_host(C.byref(params), C.byref(h)), "rtmc_prepare_material_host")
    try:
        nch = 1 if params.acoustics else 4
        tables = []
        for c in range(nch):
            n = int(h.n_knots[c])
            tables.append((np.ctypeslib.as_array(h.x[c], shape=(n,)).copy(),
                           np.ctypeslib.as_array(h.y[c], shape=(n,)).copy()))
        return {"Sigma": np.array(h.Sigma[:nch]), "Sigma_prime": np.array(h.Sigma_prime[:nch]),
                "mft": np.array(h.mft[:]), "P2P": h.P2P, "S2S": h.S2S, "tables": tables,
                "extrap": "nearest" if h.extrap == _abi.EXTRAP_NEAREST else "linear"}
    finally:
        lib.rtmc_free_material_host(C.byref(h))


class NativeMaterial:
    """rtmc_prepare_material: the tables in device memory, freed with the object."""

    def __init__(self, params):
        self._lib = _abi.load()
        self.c = _abi.Material()
        self._h = C.c_void_p()
        _abi.check(self._lib.rtmc_prepare_material(C.byref(params), C.byref(self.c), C.byref(self._h)),
                   "rtmc_prepare_material")

    def close(self):
        if self._h:
            self._lib.rtmc_free_material(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NativeMaterials:
    """rtmc_prepare_materials: n materials prepared at once (one per host
    thread), their tables in one device block freed with the object.  ``items``
    holds one object per material whose ``.c`` is its rtmc_material."""

    class _Item:
        def __init__(self, c):
            self.c = c

    def __init__(self, params_list):
        self._lib = _abi.load()
        n = len(params_list)
        arr = (_abi.MaterialParams * max(n, 1))(*params_list)
        self._mats = (_abi.Material * max(n, 1))()
        self._h = C.c_void_p()
        _abi.check(self._lib.rtmc_prepare_materials(arr, n, self._mats, C.byref(self._h)), "rtmc_prepare_materials")
        self.items = [NativeMaterials._Item(self._mats[i]) for i in range(n)]

    def close(self):
        if self._h:
            self._lib.rtmc_free_material(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def lc_derivatives(params_list, n_theta: int = 2049, rel: float = 1e-4):
    """rtmc_lc_derivatives: (g, dlogSigma) with g[i] = d log sigma / d lc of
    material i on linspace(0, pi, n_theta) and dlogSigma[i] = d log Sigma / d lc."""
    lib = _abi.load()
    n = len(params_list)
    arr = (_abi.MaterialParams * max(n, 1))(*params_list)
    g = np.zeros((n, n_theta))
    dS = np.zeros(n)
    _abi.check(lib.rtmc_lc_derivatives(arr, n, int(n_theta), float(rel),
                                       g.ctypes.data_as(C.POINTER(C.c_double)),
                                       dS.ctypes.data_as(C.POINTER(C.c_double))), "rtmc_lc_derivatives")
    return g, dS

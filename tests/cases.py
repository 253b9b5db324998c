"""Shared problem definitions for the parity tests, the smoke check and bench.

Each builder returns plain MATLAB-style structs (dicts) and a MaterialClass,
mirroring the reference's Examples/ scripts and the benchmark configurations.
"""
import math

import numpy as np

from radiativetransfermontecarlo_amd.material import MaterialClass


def iso_acoustic(d=3, v=2.0):
    """preset(1) with velocity v: isotropic scattering, Sigma = 1 (3-D) or 1/2 (2-D)."""
    m = MaterialClass.preset(1)
    m.v = v
    return m


def elastic_iso(d=3, rho=None):
    """Hand-set isotropic elastic cross-sections of IsotropicElastic{2D,3D}.m:28-33."""
    m = MaterialClass.preset(3)
    vp, vs = m.vp, m.vs
    K = vp / vs
    Spp = 0.2 * vp
    Sps = 0.2 * vp
    Sp = Spp + Sps
    Ssp = Sps / ((d - 1) * K ** d)
    Sss = Sp - Ssp
    c = 1 / (2 * math.pi ** (d / 2) / math.gamma(d / 2))
    m.sigma = [[lambda th, s=Spp: c * s * np.ones(np.shape(th)), lambda th, s=Sps: c * s * np.ones(np.shape(th))],
               [lambda th, s=Ssp: c * s * np.ones(np.shape(th)), lambda th, s=Sss: c * s * np.ones(np.shape(th))]]
    m.rho = rho
    return m


def config1(n=1_000_000):
    """3-D isotropic acoustic, unbounded, spherical frame (README quick start)."""
    geometry = {'dimension': 3, 'frame': 'spherical'}
    source = {'numberParticles': n, 'type': 'point', 'position': [0, 0, 0], 'direction': 'uniform',
              'lambda': 0.05}
    observation = {'x': np.arange(0, 6 + 1e-9, 0.05), 'y': [-np.pi, np.pi], 'z': [-np.pi / 2, np.pi / 2],
                   'directions': [0, np.pi], 'time': np.arange(0, 5 + 1e-9, 0.02)}
    material = MaterialClass(geometry, 10, True, 1, [0.05, 0.05], 0.0, 'exp', 0.1)
    material.Q = math.inf
    return geometry, source, material, observation


def config2(n=10_000_000):
    """2-D isotropic-medium elastic P-SV, cartesian frame, P point source, Q = 500;
    x, y = -50:0.5:50, 36 direction bins, t = 0:0.05:20 (three binned variables)."""
    geometry = {'dimension': 2, 'frame': 'cartesian'}
    source = {'numberParticles': n, 'type': 'point', 'position': [0, 0], 'polarization': 'P', 'lambda': 0.002}
    edges = np.arange(-50, 50 + 1e-9, 0.5)
    observation = {'x': edges, 'y': edges, 'directions': np.linspace(0, np.pi, 37),
                   'time': np.arange(0, 20 + 1e-9, 0.05)}
    material = MaterialClass(geometry, 5, False, [6, 3.46], [0.1, 0.1, 0.1], [0, 0, 0], 'gaussian', 0.5)
    material.Q = 500
    return geometry, source, material, observation


def config3(n=50_000_000):
    """3-D elastic half-space z < 0 with a free surface at z = 0, cylindrical frame,
    von Karman ACF (nu = 0.5)."""
    geometry = {'dimension': 3, 'frame': 'cylindrical', 'bnd': [{'dir': 3, 'val': 0.0}]}
    source = {'numberParticles': n, 'type': 'point', 'position': [0, 0, -2], 'polarization': 'P', 'lambda': 0.1}
    observation = {'x': np.arange(0, 30 + 1e-9, 0.1), 'directions': np.linspace(0, np.pi, 91),
                   'time': np.arange(0, 15 + 1e-9, 0.05)}
    material = MaterialClass(geometry, 20, False, [6, 6 / math.sqrt(3)], [0.05, 0.05, 0.05], [0, 0, 0],
                             'VonKarman', 0.2)
    material.SpectralParam = {'nu': 0.5}
    material.rho = 2000
    material.Q = math.inf
    return geometry, source, material, observation


def config4(n_items=64, seed=20240604):
    """64 3-D acoustic materials on the config1 geometry and observation:
    cv = [c c], c ~ U[0.02, 0.15], lc ~ U[0.05, 0.5], Q ~ U[50, 1000], exp ACF,
    f = 10, v = 1; particles per material log-uniform in [1e5, 1e7] (numpy
    default_rng(seed)).  Returns (geometry, source, base material,
    observation, params (n x 3: c, lc, Q), N (n))."""
    geometry, source, base, observation = config1()
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.02, 0.15, n_items)
    lc = rng.uniform(0.05, 0.5, n_items)
    Q = rng.uniform(50.0, 1000.0, n_items)
    N = np.round(10.0 ** rng.uniform(5.0, 7.0, n_items)).astype(np.int64)
    params = np.stack([c, lc, Q], axis=1)
    return geometry, source, base, observation, params, N


def config5(n=1_000_000_000):
    """Large 3-D acoustic, cartesian frame: exp ACF lc = 0.05, f = 40, cv = [0.1 0.1],
    Q = 300; 121^3 cells of 0.1 (edges -6.05:0.1:6.05), 8 direction bins,
    t = 0:0.02:10 (four binned variables, a 28.4 GB histogram)."""
    geometry = {'dimension': 3, 'frame': 'cartesian'}
    source = {'numberParticles': n, 'type': 'point', 'position': [0, 0, 0], 'direction': 'uniform',
              'lambda': 0.05}
    edges = np.round(np.arange(-6.05, 6.05 + 1e-9, 0.1), 10)
    observation = {'x': edges, 'y': edges, 'z': edges, 'directions': np.linspace(0, np.pi, 9),
                   'time': np.arange(0, 10 + 1e-9, 0.02)}
    material = MaterialClass(geometry, 40, True, 1, [0.1, 0.1], 0.0, 'exp', 0.05)
    material.Q = 300.0
    return geometry, source, material, observation

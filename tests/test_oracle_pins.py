"""Pins of the CPU oracle against known answers and closed forms (CPU only).

The reference (MATLAB) cannot run here and ships no numeric golden vectors,
so the oracle is pinned by: the published Philox known-answer vectors, and
closed forms / invariants that a dropped term, a wrong sign or a transposed
index would break (ballistic flight, rotation invariants, histogram edge
rules, Q-decay law, mode-conversion equilibrium).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from radiativetransfermontecarlo_amd import _abi
from radiativetransfermontecarlo_amd.material import MaterialClass, InverseCDF
from radiativetransfermontecarlo_amd.observation import initializeObservation
from radiativetransfermontecarlo_amd import solver


# Random123 kat_vectors for philox4x32-10 (Salmon et al., SC'11 software release)
KAT = [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,out", KAT)
def test_philox_known_answers(ctr, key, out):
    assert tuple(orc.philox(ctr, key)) == out


def _iso_material(d=3, v=2.0):
    m = MaterialClass.preset(1)
    m.v = v
    return MaterialClass.prepareSigma(m, d)


def test_ballistic_flight_is_exact_line():
    """mft = Inf: x(T) = x0 + v T dir for every particle, dir unchanged (up to renormalisation)."""
    m = _iso_material()
    m.meanFreeTime = math.inf
    hm = orc.HostMaterial(m)
    N = 500
    P = orc.init_particles({'lambda': 0.3}, 3, True, orc.HostDraws.philox(7), N)
    x0, d0 = P.x.copy(), P.dir.copy()
    g = solver.build_geometry({'dimension': 3})
    info = orc.propagate_smalldt(hm, g, P, 1.5, orc.HostDraws.philox(7, 1), T_eff=1.5)
    assert info.n_sub == 1 and info.dt == 1.5
    np.testing.assert_allclose(P.x, x0 + 2.0 * 1.5 * d0, rtol=0, atol=1e-13)
    np.testing.assert_allclose(P.dir, d0, atol=1e-15)
    np.testing.assert_allclose(P.t, 1.5)


@pytest.mark.parametrize("d", [2, 3])
def test_rotation_keeps_orthonormal_frame(d):
    m = _iso_material(d)
    hm = orc.HostMaterial(m)
    P = orc.init_particles({'lambda': 0.01}, d, True, orc.HostDraws.philox(3), 2000)
    g = solver.build_geometry({'dimension': d})
    for it in range(1, 6):
        orc.propagate_smalldt(hm, g, P, it * 2.0, orc.HostDraws.philox(3, it), T_eff=2.0)
    nd = np.sum(P.dir ** 2, axis=0)
    npp = np.sum(P.perp ** 2, axis=0)
    dp = np.sum(P.dir * P.perp, axis=0)
    np.testing.assert_allclose(nd, 1, atol=1e-14)
    np.testing.assert_allclose(npp, 1, atol=1e-14)
    np.testing.assert_allclose(dp, 0, atol=1e-13)
    if d == 2:  # in-plane motion: dir_z stays at the 6e-17 level of cos(pi/2)
        assert np.max(np.abs(P.dir[2])) < 1e-15


def test_point_source_birth_law():
    """|x - x0| is half-normal with scale lambda; directions uniform on the sphere."""
    N = 200_000
    P = orc.init_particles({'lambda': 0.5, 'position': [1, 2, 3]}, 3, True, orc.HostDraws.philox(11), N)
    r = np.linalg.norm(P.x - np.array([[1], [2], [3]]), axis=0)
    # E|N(0, s^2)| = s sqrt(2/pi)
    assert abs(r.mean() - 0.5 * math.sqrt(2 / math.pi)) < 4 * 0.5 * 0.6 / math.sqrt(N)
    for k in range(3):
        assert abs(P.dir[k].mean()) < 5 / math.sqrt(3 * N)
        assert abs((P.dir[k] ** 2).mean() - 1 / 3) < 5 * 0.3 / math.sqrt(N)
    np.testing.assert_allclose(np.sum(P.dir * P.perp, axis=0), 0, atol=1e-14)


def test_q_decay_law():
    """Alive fraction after time t is exp(-omega t / Q) (binomial tolerance)."""
    m = _iso_material()
    m.Q = 50.0
    m.Frequency = 1.0
    hm = orc.HostMaterial(m)
    N = 40_000
    P = orc.init_particles({'lambda': 0.1}, 3, True, orc.HostDraws.philox(5), N)
    g = solver.build_geometry({'dimension': 3})
    T = 0.0
    for it in range(1, 11):
        T += 0.7
        orc.propagate_smalldt(hm, g, P, T, orc.HostDraws.philox(5, it), T_eff=0.7)
    expect = math.exp(-2 * math.pi * 1.0 * T / 50.0)
    frac = P.alive.mean()
    assert abs(frac - expect) < 4 * math.sqrt(expect * (1 - expect) / N)


def test_histcounts_edge_rules():
    e = np.array([0.0, 1.0, 2.0, 3.0])
    lib = orc.load()
    ptr = e.ctypes.data_as(orc.C.POINTER(orc.C.c_double))
    cases = {-1e-300: -1, 0.0: 0, 0.999: 0, 1.0: 1, 2.5: 2, 3.0: 2, 3.0000001: -1, float('nan'): -1}
    for v, k in cases.items():
        assert lib.orc_bin_export(ptr, 4, v) == k, v
    # repeated edge: the empty bin [1,1) is skipped
    e2 = np.array([0.0, 1.0, 1.0, 2.0])
    p2 = e2.ctypes.data_as(orc.C.POINTER(orc.C.c_double))
    assert lib.orc_bin_export(p2, 4, 1.0) == 2


def test_observe_matches_bruteforce_histogram():
    """observeTime on random particles = numpy brute force with the histcounts rules."""
    rng = np.random.default_rng(0)
    N = 5000
    geo = {'dimension': 3, 'frame': 'spherical'}
    obsd = {'x': np.linspace(0, 2, 9), 'y': [-np.pi, np.pi], 'z': [-np.pi / 2, np.pi / 2],
            'directions': np.linspace(0, np.pi, 5), 'time': [0, 1]}
    O = initializeObservation(geo, True, obsd, N, allow_nd=False)
    assert O.ibins == [1, 4]
    x = rng.normal(size=(3, N)) * 0.6
    dd = rng.normal(size=(3, N))
    dd /= np.linalg.norm(dd, axis=0)
    alive = (rng.random(N) > 0.1).astype(np.uint8)
    P = orc.HostParticles.from_arrays(3, x, dd, dd, np.ones(N, np.uint8), alive, np.zeros(N))
    E = np.zeros(O.size, dtype=np.uint32)
    orc.observe(solver.build_geometry(geo), orc.HostBins(O), True, P, E, 1)
    r = np.sqrt(np.sum(x ** 2, axis=0))
    psi = np.clip(dd[2], -1, 1)
    ref = np.zeros((8, 4), dtype=np.int64)
    for i in range(N):
        if not alive[i]:
            continue
        kr = np.searchsorted(O.edges[0], r[i], side='right') - 1
        if r[i] == O.edges[0][-1]:
            kr = 7
        if not (0 <= kr < 8):
            continue
        kp = np.searchsorted(O.edges[3], psi[i], side='right') - 1
        if psi[i] == O.edges[3][-1]:
            kp = 3
        ref[kr, kp] += 1
    got = E.reshape(O.shape, order='F')[:, :, 1, 0]
    np.testing.assert_array_equal(got, ref)


def test_replay_vs_philox_consistency():
    """A replay buffer holding exactly the Philox variates reproduces the Philox run."""
    m = _iso_material()
    hm = orc.HostMaterial(m)
    N = 300
    P0 = orc.init_particles({'lambda': 0.2}, 3, True, orc.HostDraws.philox(9), N)
    g = solver.build_geometry({'dimension': 3})
    D = orc.HostDraws.philox(9, 4)
    info = orc.propagate_smalldt(hm, g, P0.copy(), 1.0, D, T_eff=1.0)
    ns = int(info.n_sub)
    lib = orc.load()
    vals = np.zeros((ns, 6, N))
    Dc = D.c()
    for s in range(ns):
        for slot in range(6):
            for i in range(N):
                vals[s, slot, i] = lib.orc_draw_export(orc.C.byref(Dc), N, i, slot, s)
    # replay stores the Poisson count and the 2-D bit; build them from the uniforms
    lam = info.dt / m.meanFreeTime
    vals[:, 1, :] = (vals[:, 1, :] < 1.0 - np.exp(-lam)).astype(np.float64)
    Pa = P0.copy()
    Pb = P0.copy()
    orc.propagate_smalldt(hm, g, Pa, 1.0, D, T_eff=1.0)
    orc.propagate_smalldt(hm, g, Pb, 1.0, orc.HostDraws.recorded(vals), T_eff=1.0)
    np.testing.assert_array_equal(Pa.x, Pb.x)
    np.testing.assert_array_equal(Pa.dir, Pb.dir)


def test_elastic_equipartition():
    """Long-time S/P energy ratio of isotropic elastic scattering -> 2 (vp/vs)^3 in 3-D
    (detailed balance of P2P/S2S with Sigma_PS / Sigma_SP = 2 K^3 ... for the
    hand-set cross-sections of IsotropicElastic3D.m:28-30)."""
    d = 3
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
    MaterialClass.prepareSigma(m, d)
    hm = orc.HostMaterial(m)
    N = 20000
    P = orc.init_particles({'lambda': 0.01}, d, False, orc.HostDraws.philox(21), N)
    g = solver.build_geometry({'dimension': d})
    T = 0.0
    for it in range(1, 9):
        T += 1.0
        orc.propagate_smalldt(hm, g, P, T, orc.HostDraws.philox(21, it), T_eff=1.0)
    frac_s = 1 - P.p.mean()
    # stationary: n_P * rate(P->S) = n_S * rate(S->P)
    rPS = m.Sigma[0, 1]
    rSP = m.Sigma[1, 0]
    expect = rPS / (rPS + rSP)
    assert abs(frac_s - expect) < 5 * math.sqrt(expect * (1 - expect) / N)


@pytest.mark.parametrize("d", [3, 2])
def test_observe_position_angle_bruteforce(d):
    """direction_variable='position' (README.md:209): the direction bin is the
    cosine (3-D) or the angle (2-D) between the direction and the position
    vector; numpy brute force, and a particle at the origin is not counted."""
    rng = np.random.default_rng(4)
    N = 4000
    geo = {'dimension': d, 'frame': 'spherical'}
    obsd = {'x': np.linspace(0, 2, 9), 'directions': np.linspace(0, np.pi, 7), 'time': [0, 1],
            'direction_variable': 'position'}
    O = initializeObservation(geo, True, obsd, N)
    assert O.ibins == [1, 4] and O.direction_variable == 'position'
    x = rng.normal(size=(3, N)) * 0.6
    dd = rng.normal(size=(3, N))
    if d == 2:
        x[2] = 0.0
        dd[2] = 0.0
    dd /= np.linalg.norm(dd, axis=0)
    x[:, 0] = 0.0  # at the origin: no angle
    P = orc.HostParticles.from_arrays(d, x, dd, dd, np.ones(N, np.uint8), np.ones(N, np.uint8), np.zeros(N))
    E = np.zeros(O.size, dtype=np.uint32)
    orc.observe(solver.build_geometry(geo), orc.HostBins(O), True, P, E, 1)
    r = np.sqrt(np.sum(x ** 2, axis=0))
    with np.errstate(invalid='ignore'):
        c = np.clip(np.sum(dd * x, axis=0) / r, -1, 1)
    v = c if d == 3 else np.arccos(c)
    nb = O.edges[3].size - 1
    ref = np.zeros((8, nb), dtype=np.int64)
    for i in range(N):
        if not np.isfinite(v[i]):
            continue
        kr = min(np.searchsorted(O.edges[0], r[i], side='right') - 1, 7)
        if not (0 <= kr < 8) or r[i] > O.edges[0][-1]:
            continue
        kp = min(np.searchsorted(O.edges[3], v[i], side='right') - 1, nb - 1)
        ref[kr, kp] += 1
    got = E.reshape(O.shape, order='F')[:, :, 1, 0]
    assert ref.sum() == np.sum(np.isfinite(v) & (r <= 2)) and not np.isfinite(v[0])
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("d", [3, 2])
def test_outgoing_ballistic_source_position_angle_closed_form(d):
    """An 'outgoing' point source (direction = position vector) in a medium that
    does not scatter within the window: every particle keeps cos = 1 (angle 0),
    so with direction_variable='position' every deposit is in the direction bin
    that holds it, N per observation time."""
    geometry = {'dimension': d, 'frame': 'spherical'}
    source = {'lambda': 0.2, 'direction': 'outgoing', 'position': [0.0] * d}
    m = MaterialClass(geometry, 10, True, 1, [1e-7, 1e-7], 0.0, 'exp', 0.1)
    m = MaterialClass.prepareSigma(m, d)
    assert float(np.asarray(m.meanFreeTime).ravel()[0]) > 1e6
    observation = {'x': np.linspace(0, 3, 31), 'directions': np.linspace(0, np.pi, 5),
                   'time': np.arange(0, 0.5 + 1e-9, 0.05), 'direction_variable': 'position'}
    N = 20_000
    O = initializeObservation(geometry, True, observation, N)
    E, nd = orc.forward(geometry, source, m, O, N, seed=3)
    H = E.reshape(O.shape, order='F')
    top = H.shape[1] - 1 if d == 3 else 0   # cos = 1 is the last cos-bin (3-D); angle 0 the first (2-D)
    per_t = H.sum(axis=(0, 1, 3))
    assert np.all(per_t[1:] == N)
    assert H[:, top, 1:, :].sum() == N * (O.Nt - 1)

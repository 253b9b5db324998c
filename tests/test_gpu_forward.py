"""The fused rtmc_forward kernel against the CPU oracle (statistical parity).

The fused kernel samples the sub-step process of propagateParticleSmallDt
event to event with its own Philox streams; the oracle runs the reference's
sub-step loop literally.  Different random streams: the histograms must agree
bin by bin within Monte Carlo error (tests/stats.py).  Exact invariants:
with Q = Inf and bins covering every reachable position, obs.energy(t > 0)
is exactly 1.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from radiativetransfermontecarlo_amd import _abi, solver
from radiativetransfermontecarlo_amd.material import MaterialClass
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import cases
from tests.stats import assert_same_distribution

pytestmark = pytest.mark.gpu


def _run_both(geometry, source, material, observation, n_gpu, n_cpu, time_steps=0, seed=11, hist_mode=0,
              allow_nd=False):
    d = geometry['dimension']
    material = MaterialClass.prepareSigma(material, d)
    material.timeSteps = time_steps
    O = initializeObservation(geometry, bool(material.acoustics), observation, n_gpu, allow_nd=allow_nd)
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, n_gpu, seed=seed, device='cuda',
                                          time_steps=time_steps, hist_mode=hist_mode)
    Eg = E.cpu().numpy().view(np.uint32)
    Eo, _ = orc.forward(geometry, source, material, O, n_cpu, seed=seed + 1, time_steps=time_steps,
                        packet=10_000)
    return O, Eg, Eo, cnt.cpu().numpy()


@pytest.mark.parametrize("hist_mode", [0, 1])
def test_config1_shape_statistics(gpu, hist_mode):
    geometry, source, material, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 2 + 1e-9, 0.02))
    O, Eg, Eo, cnt = _run_both(geometry, source, material, observation, 1_000_000, 100_000, hist_mode=hist_mode)
    # every particle is inside r < 6 up to t = 2: each time slice holds all of them
    per_t = Eg.reshape(O.shape, order='F').sum(axis=(0, 1, 3))
    assert np.all(per_t[1:] == 1_000_000) and per_t[0] == 0
    assert cnt[_abi.CNT_DEPOSIT] == 1_000_000 * (O.Nt - 1)
    assert_same_distribution(Eg, 1_000_000, Eo, 100_000, "config1")


def test_elastic_2d_q_statistics(gpu):
    geometry = {'dimension': 2, 'frame': 'cartesian'}
    source = {'numberParticles': 1, 'lambda': 0.01, 'position': [0, 0]}
    e = np.linspace(-3, 3, 25)
    observation = {'x': e, 'y': e, 'time': np.arange(0, 0.8 + 1e-9, 0.05)}
    m = cases.elastic_iso(2)
    m.Q = [40.0, 20.0]
    m.Frequency = 1.0
    O, Eg, Eo, cnt = _run_both(geometry, source, m, observation, 1_000_000, 60_000)
    assert cnt[_abi.CNT_KILLED] > 0
    assert_same_distribution(Eg, 1_000_000, Eo, 60_000, "elastic 2-D with Q")


def test_elastic_half_space_statistics(gpu):
    """Free surface z = 0 (MaterialInterface + Zoeppritz), cylindrical frame."""
    geometry = {'dimension': 3, 'frame': 'cylindrical', 'bnd': [{'dir': 3, 'val': 0.0}]}
    source = {'lambda': 0.05, 'position': [0, 0, -0.3]}
    observation = {'x': np.linspace(0, 4, 21), 'z': np.linspace(-4, 0, 9), 'time': np.arange(0, 0.6 + 1e-9, 0.05)}
    m = cases.elastic_iso(3, rho=2000)
    O, Eg, Eo, cnt = _run_both(geometry, source, m, observation, 1_000_000, 40_000)
    assert cnt[_abi.CNT_BOUNDARY] > 0
    # nothing ever sits above the free surface
    assert_same_distribution(Eg, 1_000_000, Eo, 40_000, "elastic half-space")


def test_acoustic_cylinder_and_absorbing_statistics(gpu):
    geometry = {'dimension': 3, 'frame': 'cylindrical',
                'bnd': [{'dir': 4, 'val': 1.0}, {'dir': 3, 'val': 1.5, 'type': 'absorbing'}]}
    source = {'lambda': 0.05}
    observation = {'x': np.linspace(0, 1.2, 13), 'z': np.linspace(-3, 3, 13), 'time': np.arange(0, 2 + 1e-9, 0.1)}
    m = cases.iso_acoustic(3, v=2.0)
    O, Eg, Eo, cnt = _run_both(geometry, source, m, observation, 1_000_000, 60_000)
    assert cnt[_abi.CNT_KILLED] > 0
    # the cylinder keeps every particle at r <= 1
    Eh = Eg.reshape(O.shape, order='F')
    assert Eh[10:].sum() == 0
    assert_same_distribution(Eg, 1_000_000, Eo, 60_000, "cylinder + absorbing plane")


@pytest.mark.parametrize("d", [2, 3])
def test_event_driven_statistics(gpu, d):
    """material.timeSteps = 1: exponential free flight (propagateParticle.m)."""
    geometry = {'dimension': d, 'frame': 'spherical'}
    source = {'lambda': 0.02}
    observation = {'x': np.linspace(0, 4, 41), 'time': np.arange(0, 1.5 + 1e-9, 0.05)}
    m = MaterialClass({'dimension': d}, 10, True, 2, [0.1, 0.1], 0.0, 'exp', 0.1)
    O, Eg, Eo, cnt = _run_both(geometry, source, m, observation, 1_000_000, 60_000, time_steps=1)
    assert cnt[_abi.CNT_SCATTER] > 0
    assert_same_distribution(Eg, 1_000_000, Eo, 60_000, f"event-driven {d}-D")


def test_three_binned_variables(gpu):
    """N-D histogram (x, y, directions) beyond the reference's two variables: its
    2-D marginal equals the reference-compatible (x, y) histogram in law."""
    geometry = {'dimension': 2, 'frame': 'cartesian'}
    source = {'lambda': 0.01, 'position': [0, 0]}
    e = np.linspace(-2, 2, 17)
    m = cases.elastic_iso(2)
    obs3 = {'x': e, 'y': e, 'directions': np.linspace(0, np.pi, 5), 'time': np.arange(0, 0.5 + 1e-9, 0.05)}
    O3, Eg3, _, _ = _run_both(geometry, source, m, obs3, 500_000, 1000, allow_nd=True)
    obs2 = dict(obs3, directions=[0, np.pi])
    O2, Eg2, Eo2, _ = _run_both(geometry, source, m, obs2, 500_000, 50_000, seed=5)
    marg = Eg3.reshape(O3.shape, order='F').sum(axis=2)
    assert marg.sum() == Eg3.sum()
    assert_same_distribution(marg.reshape(-1, order='F'), 500_000, Eo2, 50_000, "3-variable marginal")


def test_radiative_transfer_dropin_fields(gpu):
    geometry, source, material, observation = cases.config1(n=150_000)
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.05))
    obs = solver.radiativeTransfer(geometry, source, material, observation, seed=3)
    for k in ('d', 'acoustics', 'N', 't', 'Nt', 'x', 'Nx', 'dx', 'binX', 'y', 'Ny', 'dy', 'binY', 'z', 'Nz', 'dz',
              'binZ', 'psi', 'Npsi', 'dpsi', 'binPsi', 'energyDensity', 'energy'):
        assert k in obs, k
    assert obs['N'] == 200_000                                   # packets of 1e5 (radiativeTransfer.m:23-28)
    assert obs['energyDensity'].shape == (120, 1, obs['Nt'])
    assert obs['energy'][0] == 0.0
    np.testing.assert_allclose(obs['energy'][1:], 1.0, rtol=0, atol=0)
    # integral of the density over the spherical shells recovers the energy
    integ = (obs['energyDensity'][:, 0, :] * obs['dx'][:, None]).sum(axis=0) * obs['dy'][0] * obs['dz'][0]
    np.testing.assert_allclose(integ, obs['energy'], rtol=1e-12)


def test_forward_chunks_add_up(gpu):
    """Two calls on disjoint particle ranges equal one call on the union (exactly:
    the Philox stream of a particle depends only on its global id)."""
    geometry, source, material, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 0.6 + 1e-9, 0.02))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1)
    E1, _, _ = solver.forward_histogram(geometry, source, material, O, 300_000, seed=9, device='cuda')
    E2, _, _ = solver.forward_histogram(geometry, source, material, O, 100_000, seed=9, device='cuda')
    solver.forward_histogram(geometry, source, material, O, 200_000, seed=9, device='cuda', particle_offset=100_000,
                             E=E2)
    assert torch.equal(E1, E2)


def test_paasschens_2d_exact_solution(gpu):
    """Event-driven 2-D isotropic acoustic scattering against the exact 2-D
    solution of Paasschens (1997) (+Comparison/analyticalPaasschens.m:65-68),
    inside the ballistic front, within 3 % (bin averaging + Monte Carlo error)."""
    from tests.analytic import paasschens_2d_diffuse
    geometry = {'dimension': 2, 'frame': 'spherical'}
    source = {'lambda': 1e-4}
    observation = {'x': np.arange(0, 10 + 1e-9, 0.1), 'time': np.arange(0, 4 + 1e-9, 0.1)}
    m = MaterialClass.prepareSigma(cases.iso_acoustic(2, v=2.0), 2)
    m.timeSteps = 1
    N = 10_000_000
    O = initializeObservation(geometry, True, observation, N)
    E, _, _ = solver.forward_histogram(geometry, source, m, O, N, seed=17, device='cuda', time_steps=1)
    from radiativetransfermontecarlo_amd.observation import normalize
    dens, energy = normalize(O, E.cpu().numpy().view(np.uint32), N)
    np.testing.assert_array_equal(energy[1:], 1.0)
    r = O.obs['x']
    for it in (20, 30, 40):
        t = O.t[it]
        ref = paasschens_2d_diffuse(r, t, m.Sigma, 2.0)
        sel = (r > 0.5) & (r < 0.7 * 2.0 * t)
        rel = dens[sel, 0, it] / ref[sel] - 1
        assert np.max(np.abs(rel)) < 0.03, (t, np.max(np.abs(rel)))


@pytest.mark.parametrize("case", ["config1", "elastic2d", "halfspace"])
def test_fp32_flight_statistics(gpu, case):
    """RTMC_PREC_FP32_FLIGHT: FP32 direction/position arithmetic, FP64 time.
    Per-bin agreement with the FP64 oracle within Monte Carlo error (the FP32
    rounding of positions, ~1e-7 relative, is far below the bin widths)."""
    if case == "config1":
        geometry, source, material, observation = cases.config1()
        observation = dict(observation, time=np.arange(0, 2 + 1e-9, 0.02))
        n_cpu = 100_000
    elif case == "elastic2d":
        geometry = {'dimension': 2, 'frame': 'cartesian'}
        source = {'lambda': 0.01, 'position': [0, 0]}
        e = np.linspace(-3, 3, 25)
        observation = {'x': e, 'y': e, 'time': np.arange(0, 0.8 + 1e-9, 0.05)}
        material = cases.elastic_iso(2)
        material.Q = [40.0, 20.0]
        material.Frequency = 1.0
        n_cpu = 60_000
    else:
        geometry = {'dimension': 3, 'frame': 'cylindrical', 'bnd': [{'dir': 3, 'val': 0.0}]}
        source = {'lambda': 0.05, 'position': [0, 0, -0.3]}
        observation = {'x': np.linspace(0, 4, 21), 'z': np.linspace(-4, 0, 9),
                       'time': np.arange(0, 0.6 + 1e-9, 0.05)}
        material = cases.elastic_iso(3, rho=2000)
        n_cpu = 40_000
    d = geometry['dimension']
    material = MaterialClass.prepareSigma(material, d)
    O = initializeObservation(geometry, bool(material.acoustics), observation, 1)
    E32, cnt, _ = solver.forward_histogram(geometry, source, material, O, 1_000_000, seed=21, device='cuda',
                                           precision='fp32')
    Eg = E32.cpu().numpy().view(np.uint32)
    if case == "config1":
        per_t = Eg.reshape(O.shape, order='F').sum(axis=(0, 1, 3))
        assert np.all(per_t[1:] == 1_000_000)
    Eo, _ = orc.forward(geometry, source, material, O, n_cpu, seed=22, packet=10_000)
    assert_same_distribution(Eg, 1_000_000, Eo, n_cpu, f"fp32 flight {case}")


def test_fp32_flight_per_bin_error_against_fp64(gpu):
    """The FP32 free-flight variant against the FP64 kernel on a scaled-down
    version of the 1e9-particle workload (3-D cartesian, four binned variables,
    Q = 300, time-major histogram): 1e7 particles each, independent streams.
    Every bin agrees within Monte Carlo error, and the per-time energy within
    its own standard error."""
    geometry, source, material, observation = cases.config5(n=1)
    e = np.round(np.arange(-3.1, 3.1 + 1e-9, 0.2), 10)
    observation = dict(observation, x=e, y=e, z=e, time=np.arange(0, 2 + 1e-9, 0.02))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1, allow_nd=True)
    n = 10_000_000
    E64, c64, _k = solver.forward_histogram(geometry, source, material, O, n, seed=41, device='cuda')
    E32, c32, _k = solver.forward_histogram(geometry, source, material, O, n, seed=42, device='cuda',
                                            precision='fp32')
    a = E64.cpu().numpy().view(np.uint32)
    b = E32.cpu().numpy().view(np.uint32)
    st = assert_same_distribution(a, n, b, n, "fp32 vs fp64")
    per_t64 = a.reshape(O.shape, order='F').sum(axis=(0, 1, 2, 3)).ravel().astype(np.float64)
    per_t32 = b.reshape(O.shape, order='F').sum(axis=(0, 1, 2, 3)).ravel().astype(np.float64)
    sel = per_t64 > 0
    rel = np.abs(per_t32[sel] - per_t64[sel]) / per_t64[sel]
    se = np.sqrt(2.0 / per_t64[sel])
    print(f"fp32 vs fp64: {st}; per-time energy max rel diff {rel.max():.2e} (se {se.max():.2e})")
    assert np.all(rel < 5 * se + 1e-6)


@pytest.mark.parametrize("frame,d", [("spherical", 3), ("cylindrical", 3), ("spherical", 2)])
def test_radial_bins_fine_global_edges_match_coarse(gpu, frame, d):
    """Radius-only binning: 5000 radial bins (edges too many for LDS, read from
    global memory, bin from the radius) summed 50 at a time equal 100 bins (edges
    in LDS, bin from the squared radius) of the same trajectories (same seed),
    up to positions within rounding of an edge."""
    geometry = {'dimension': d, 'frame': frame}
    source = {'numberParticles': 1, 'type': 'point', 'position': [0.0] * d, 'lambda': 0.05}
    material = MaterialClass.prepareSigma(cases.iso_acoustic(d), d)
    t = np.arange(0, 1.0 + 1e-9, 0.05)
    n = 200_000

    def hist(nx):
        obs = {'x': np.linspace(0, 3, nx + 1), 'time': t}
        if d == 3 and frame == 'cylindrical':
            obs['z'] = [-np.inf, np.inf]
        O = initializeObservation(geometry, True, obs, n)
        E, _, _ = solver.forward_histogram(geometry, source, material, O, n, seed=21, device='cuda')
        return O, E.cpu().numpy().view(np.uint32).astype(np.int64)

    Of, Ef = hist(5000)
    Oc, Ec = hist(100)
    fine = Ef.reshape(Of.shape, order='F')
    coarse = Ec.reshape(Oc.shape, order='F')
    assert fine.shape[0] == 5000 and coarse.shape[0] == 100
    summed = fine.reshape(100, 50, *fine.shape[1:]).sum(axis=1)
    assert summed.sum() == coarse.sum() > 0
    assert np.abs(summed - coarse).sum() <= 2e-5 * coarse.sum()


def test_radial_source_statistics(gpu):
    """A user radial source law (source.radial, initializeParticle.m:22-27) in
    the tracing kernel's birth, against the oracle's births and flights."""
    class Law:
        Rmax = 1.0

        def __call__(self, r):
            return 1.0 + 0.0 * r  # uniform density in r^(d-1) dr: a uniform ball

    geometry = {'dimension': 3, 'frame': 'spherical'}
    source = {'lambda': 0.2, 'radial': Law()}
    material = cases.iso_acoustic(3)
    observation = {'x': np.linspace(0, 2, 21), 'time': np.arange(0, 0.3 + 1e-9, 0.05)}
    O, Eg, Eo, cnt = _run_both(geometry, source, material, observation, 500_000, 50_000)
    assert_same_distribution(Eg, 500_000, Eo, 50_000, "radial source")
    # at t = 0.05 (v = 2) every particle lies within 1 + 0.1 of the origin
    first = Eg.reshape(O.shape, order='F')[:, 0, 1, 0]
    assert first[11:].sum() == 0 and first.sum() == 500_000


@pytest.mark.parametrize("d", [3, 2])
def test_position_angle_direction_variable_statistics(gpu, d):
    """direction_variable='position' (README.md:209) in the tracing kernel: the
    direction bin follows the particle's position at every deposit; against the
    oracle in law."""
    geometry = {'dimension': d, 'frame': 'spherical'}
    source = {'lambda': 0.05, 'position': [0.0] * d}
    material = cases.iso_acoustic(d)
    observation = {'x': np.linspace(0, 2.5, 11), 'directions': np.linspace(0, np.pi, 7),
                   'time': np.arange(0, 1.0 + 1e-9, 0.1), 'direction_variable': 'position'}
    O, Eg, Eo, cnt = _run_both(geometry, source, material, observation, 500_000, 50_000)
    assert O.direction_variable == 'position'
    assert_same_distribution(Eg, 500_000, Eo, 50_000, f"position angle {d}-D")


@pytest.mark.parametrize("d", [3, 2])
def test_outgoing_ballistic_position_angle_closed_form(gpu, d):
    """An 'outgoing' source in a medium that does not scatter within the window:
    the direction stays along the position vector, so every deposit is in the
    bin of cos = 1 (3-D) or angle 0 (2-D)."""
    geometry = {'dimension': d, 'frame': 'spherical'}
    source = {'lambda': 0.2, 'direction': 'outgoing', 'position': [0.0] * d}
    m = MaterialClass.prepareSigma(MaterialClass(geometry, 10, True, 1, [1e-7, 1e-7], 0.0, 'exp', 0.1), d)
    observation = {'x': np.linspace(0, 3, 31), 'directions': np.linspace(0, np.pi, 5),
                   'time': np.arange(0, 0.5 + 1e-9, 0.05), 'direction_variable': 'position'}
    N = 300_000
    O = initializeObservation(geometry, True, observation, N)
    E, cnt, _k = solver.forward_histogram(geometry, source, m, O, N, seed=8, device='cuda')
    H = E.cpu().numpy().view(np.uint32).reshape(O.shape, order='F')
    top = H.shape[1] - 1 if d == 3 else 0
    assert np.all(H.sum(axis=(0, 1, 3))[1:] == N)
    assert H[:, top, 1:, :].sum() == N * (O.Nt - 1)

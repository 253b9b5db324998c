"""Scattering angles drawn from compressed tables held in LDS
(rtmc_compress_table, rtmc_material.lds).

- Against the oracle (the reference's full tables, its sub-step loop): every
  bin within Monte Carlo error, for the README case (LDS histogram), the 2-D
  elastic workload's own material (time-major scratch, four tables), the
  elastic half-space with its own von Karman material (free surface), and
  the event-driven process.
- Against the kernel on the full tables: the per-bin error of the compressed
  tables at the 1e9-particle workload's material (1e7 particles each).
- The tables' LDS block sits beside the histogram, the bin edges and the
  hot-counter cache: every placement gives the same histogram bit for bit.
- A batch of materials whose tables have different knot counts gives each
  item the single-problem histogram; a batch in which one material has no
  compressed tables draws every item from its full tables.
- The adjoint replays the forward's draws from the same LDS tables.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from radiativetransfermontecarlo_amd import _abi, solver
from radiativetransfermontecarlo_amd import batch as rb
from radiativetransfermontecarlo_amd.material import MaterialClass
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import cases
from tests.stats import assert_same_distribution

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _forward(geometry, source, material, O, n, seed, lds_tol=TOL, lds_max_knots=1024, **kw):
    d = geometry['dimension']
    dm = solver.DeviceMaterial(material, d, 'cuda', with_zoeppritz=len(geometry.get('bnd', [])) > 0,
                               lds_tol=lds_tol, lds_max_knots=lds_max_knots)
    E, cnt, keep = solver.forward_histogram(geometry, source, material, O, n, seed=seed, device='cuda', dmat=dm, **kw)
    return E, cnt, dm


def _oracle_case(name):
    if name == "readme":
        geometry, source, material, observation = cases.config1()
        observation = dict(observation, time=np.arange(0, 2 + 1e-9, 0.02))
        return geometry, source, material, observation, 0, 1_000_000, 100_000, {}
    if name == "elastic2d":
        geometry, source, material, observation = cases.config2()
        e = np.arange(-6, 6 + 1e-9, 0.5)
        observation = {'x': e, 'y': e, 'time': np.arange(0, 1.5 + 1e-9, 0.05)}
        return geometry, source, material, observation, 0, 1_000_000, 40_000, {'hist_mode': 3}
    if name == "halfspace":
        geometry, source, material, observation = cases.config3()
        observation = {'x': np.arange(0, 3 + 1e-9, 0.2), 'directions': np.linspace(0, np.pi, 7),
                       'time': np.arange(0, 0.5 + 1e-9, 0.05)}
        source = dict(source, position=[0, 0, -0.5])
        return geometry, source, material, observation, 0, 1_000_000, 20_000, {}
    # event-driven 3-D
    geometry = {'dimension': 3, 'frame': 'spherical'}
    source = {'lambda': 0.02}
    observation = {'x': np.linspace(0, 4, 41), 'time': np.arange(0, 1.5 + 1e-9, 0.05)}
    material = MaterialClass({'dimension': 3}, 10, True, 2, [0.1, 0.1], 0.0, 'exp', 0.1)
    return geometry, source, material, observation, 1, 1_000_000, 60_000, {}


@pytest.mark.parametrize("name", ["readme", "elastic2d", "halfspace", "event3d"])
def test_lds_table_statistics_against_oracle(gpu, name):
    geometry, source, material, observation, ts, n_gpu, n_cpu, kw = _oracle_case(name)
    d = geometry['dimension']
    material = MaterialClass.prepareSigma(material, d)
    material.timeSteps = ts
    O = initializeObservation(geometry, bool(material.acoustics), observation, n_gpu, allow_nd=True)
    E, cnt, dm = _forward(geometry, source, material, O, n_gpu, 11, time_steps=ts, **kw)
    assert all(dm.c.lds[k].m > 0 for k in range(1 if material.acoustics else 4))
    Eg = E.cpu().numpy().view(np.uint32)
    Eo, _ = orc.forward(geometry, source, material, O, n_cpu, seed=12, time_steps=ts, packet=10_000)
    st = assert_same_distribution(Eg, n_gpu, Eo, n_cpu, f"LDS tables, {name}")
    print(f"{name}: {st}, knots {[dm.c.lds[k].m for k in range(4)]}")
    if name == "halfspace":
        assert int(cnt[_abi.CNT_BOUNDARY]) > 0


def test_lds_table_per_bin_error_against_full_table(gpu):
    """The 1e9-particle workload's material (exp ACF lc = 0.05, f = 40, Q = 300)
    on a 31^3 x 8 x 101 grid, 1e7 particles each, independent streams: the
    compressed tables (tol 1e-5 rad) against the full tables."""
    geometry, source, material, observation = cases.config5(n=1)
    e = np.round(np.arange(-3.1, 3.1 + 1e-9, 0.2), 10)
    observation = dict(observation, x=e, y=e, z=e, time=np.arange(0, 2 + 1e-9, 0.02))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1, allow_nd=True)
    n = 10_000_000
    Ef, cf, _k = solver.forward_histogram(geometry, source, material, O, n, seed=51, device='cuda')
    El, cl, dm = _forward(geometry, source, material, O, n, 52)
    a = Ef.cpu().numpy().view(np.uint32)
    b = El.cpu().numpy().view(np.uint32)
    st = assert_same_distribution(a, n, b, n, "LDS tables vs full tables")
    # scatterings per particle agree within their Poisson error
    sf, sl = float(cf[_abi.CNT_SCATTER]), float(cl[_abi.CNT_SCATTER])
    assert abs(sf - sl) < 5 * math.sqrt(sf + sl)
    print(f"LDS vs full tables: {st}; knots {dm.c.lds[0].m}, tol {dm.c.lds[0].tol}")


@pytest.mark.parametrize("dim,acoustic", [(3, True), (2, False), (3, False)])
def test_lds_table_placements_equal(gpu, dim, acoustic):
    """The LDS histogram, the time-major scratch with its hot-counter cache and
    direct atomics give the same histogram with the tables in LDS."""
    geometry = {'dimension': dim, 'frame': 'cartesian'}
    if not acoustic:
        geometry['bnd'] = [{'dir': dim, 'val': 0.6}]
    source = {'numberParticles': 1, 'lambda': 0.01, 'position': [0.0] * dim}
    e = np.linspace(-2, 2, 17 if dim == 2 else 9)
    observation = {'x': e, 'y': e, 'time': np.arange(0, 0.7 + 1e-9, 0.05)}
    if dim == 3:
        observation['z'] = e
    if acoustic:
        material = MaterialClass(geometry, 10, True, 2.0, [0.1, 0.1], 0.0, 'exp', 0.2)
    else:
        material = MaterialClass(geometry, 2, False, [6, 3.46], [0.1, 0.1, 0.1], [0, 0, 0], 'gaussian', 0.5)
        material.rho = 2000
    material = MaterialClass.prepareSigma(material, dim)
    O = initializeObservation(geometry, acoustic, observation, 1, allow_nd=True)
    out = {}
    for hm in (1, 2, 3):
        E, cnt, dm = _forward(geometry, source, material, O, 60_000, 9, lds_max_knots=256, hist_mode=hm)
        out[hm] = (E.clone(), cnt.clone())
    assert int(out[1][0].sum()) > 0
    for hm in (2, 3):
        assert torch.equal(out[hm][0], out[1][0]), hm
        assert torch.equal(out[hm][1], out[1][1]), hm


def test_lds_table_batch_items_equal_single_forward(gpu):
    """Materials with different knot counts (the launch's LDS block takes the
    largest), and a batch in which one material has no compressed tables
    (every item then draws from its full tables)."""
    geometry, source, base, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.02))
    O = initializeObservation(geometry, True, observation, 1)
    params = [(0.05, 0.1, math.inf), (0.12, 0.4, 200.0), (0.03, 0.05, 500.0)]
    mats = [rb.prepare_material(base, 3, c=c, lc=lc, Q=Q) for c, lc, Q in params]
    Ns, seeds = [100_000, 37_001, 64_000], [3, 4, 5]
    B = rb.Batch(geometry, O, mats, source, Ns, seeds, device='cuda', lds_tol=TOL)
    ms = [B.dmats[i].c.lds[0].m for i in range(3)]
    assert len(set(ms)) > 1
    E, cnt = B.forward()
    for i, m in enumerate(mats):
        Es, cs, _dm = _forward(geometry, source, m, O, Ns[i], seeds[i])
        assert torch.equal(E[i], Es) and torch.equal(cnt[i], cs), f"item {i}"
    # item 1 without compressed tables: the launch uses the full tables for all
    dmats = list(B.dmats)
    dmats[1] = solver.DeviceMaterial(mats[1], 3, 'cuda')
    B2 = rb.Batch(geometry, O, dmats, source, Ns, seeds, device='cuda')
    E2, c2 = B2.forward()
    for i, m in enumerate(mats):
        Es, cs, _k = solver.forward_histogram(geometry, source, m, O, Ns[i], seed=seeds[i], device='cuda')
        assert torch.equal(E2[i], Es) and torch.equal(c2[i], cs), f"mixed batch item {i}"


def test_lds_table_switch_off(gpu, monkeypatch):
    """RTMC_LDS_TABLE=0 draws from the full tables even when compressed ones are given."""
    geometry, source, material, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.02))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1)
    Ef, cf, _k = solver.forward_histogram(geometry, source, material, O, 50_000, seed=3, device='cuda')
    monkeypatch.setenv("RTMC_LDS_TABLE", "0")
    El, cl, _dm = _forward(geometry, source, material, O, 50_000, 3)
    assert torch.equal(Ef, El) and torch.equal(cf, cl)
    monkeypatch.delenv("RTMC_LDS_TABLE")
    El2, cl2, _dm = _forward(geometry, source, material, O, 50_000, 3)
    assert not torch.equal(Ef, El2)  # the compressed tables do change the draws


@pytest.mark.parametrize("hist_mode", [1, 2])
def test_lds_table_adjoint_replays_forward(gpu, hist_mode):
    """Native tables without knot truncation (knot_threshold = 0, nearest
    extrapolation) compressed at preparation: the adjoint makes the forward's
    deposits and scatterings."""
    geometry, source, base, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.02))
    O = initializeObservation(geometry, True, observation, 1)
    params = np.array([(0.08, 0.2, 300.0), (0.05, 0.35, math.inf)])
    B = rb.Batch.from_params(geometry, O, base, params, source, [60_000, 41_000], [8, 9], device='cuda',
                             lds_tol=TOL)
    assert all(B.dmats[i].c.lds[0].m > 0 and B.dmats[i].c.lds[0].extrap == _abi.EXTRAP_NEAREST for i in range(2))
    B.configure(hist_mode=hist_mode)
    E, cnt = B.forward()
    cnt_f = cnt.clone()
    gen = torch.Generator().manual_seed(2)
    w = torch.rand((2, O.size), generator=gen, dtype=torch.float32) * 2 - 1
    B.counters.zero_()
    g = B.adjoint(w).cpu().numpy()
    assert torch.equal(B.counters, cnt_f)
    assert np.all(np.isfinite(g))


def test_lds_table_dropin_radiative_transfer(gpu):
    """radiativeTransfer(..., lds_tol=...): the reference's obs fields, every
    particle counted at every time (Q = Inf, bins covering every reachable
    radius), and the energy density integrating to the energy."""
    geometry, source, material, observation = cases.config1(n=150_000)
    observation = dict(observation, time=np.arange(0, 1 + 1e-9, 0.05))
    obs = solver.radiativeTransfer(geometry, source, material, observation, seed=3, lds_tol=TOL)
    assert obs['N'] == 200_000
    assert obs['energy'][0] == 0.0
    np.testing.assert_allclose(obs['energy'][1:], 1.0, rtol=0, atol=0)
    integ = (obs['energyDensity'][:, 0, :] * obs['dx'][:, None]).sum(axis=0) * obs['dy'][0] * obs['dz'][0]
    np.testing.assert_allclose(integ, obs['energy'], rtol=1e-12)
    with pytest.raises(ValueError):
        solver.radiativeTransfer(geometry, source, material, observation, seed=3, lds_tol=TOL, engine="functions")

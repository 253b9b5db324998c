"""The benchmark workloads' own materials and sizes on the GPU.

1. Each workload's own medium against the oracle on its full bin grid, for a
   shortened time window at an oracle-feasible particle count: the 2-D
   elastic Gaussian medium with Q = 500 (200 x 200 cells x 36 directions),
   the 3-D elastic von Karman half-space with its free surface (300 radii x
   90 directions), the 3-D acoustic exponential medium f = 40, lc = 0.05,
   Q = 300 (121^3 cells x 8 directions), the last in FP64 and FP32 flight.
   The oracle's tables come from its own restatement of the scattering model
   (oracle/sigma.py), not from the product's preparation.  Per-bin two-sample
   statistics as in tests/stats.py.
2. The 5e7-particle half-space and the 1e9-particle 28.4 GB workload at full
   size: every deposit is in the histogram, the time-major placement gives
   the counts of direct atomics bit for bit, nothing is observed above the
   free surface, and (1e9 particles) the alive fraction follows
   exp(-omega t / Q) exactly in law while no particle can have left the grid.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as orc
from oracle import sigma as osg
from radiativetransfermontecarlo_amd import _abi, solver
from radiativetransfermontecarlo_amd.material import MaterialClass
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import cases
from tests.stats import assert_same_distribution

pytestmark = pytest.mark.gpu


def _compare_populated(Eg, ng, Eo, no, what):
    """Two-sample comparison over the bins either side populates (selected on
    the integer arrays first: the 1e9-particle grid has 3e8 bins), for bins
    where the oracle's smaller sample expects at least 20 counts."""
    sel = np.flatnonzero((Eg != 0) | (Eo != 0))
    res = assert_same_distribution(Eg[sel], ng, Eo[sel], no, what, min_bins=200, min_small=20.0)
    print(f"{what}: {res}", flush=True)
    return res


def _gpu_hist(geometry, source, material, O, n, seed, precision='fp64', hist_mode=0):
    d = geometry['dimension']
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, n, seed=seed, device='cuda',
                                          precision=precision, hist_mode=hist_mode)
    torch.cuda.synchronize()
    return E, cnt


def test_config2_material_against_oracle(gpu):
    geometry, source, material, observation = cases.config2(n=1)
    observation = dict(observation, time=np.arange(0, 2 + 1e-9, 0.05))
    material = MaterialClass.prepareSigma(material, 2)
    O = initializeObservation(geometry, False, observation, 1, allow_nd=True)
    print("config 2 material: GPU 1e7, oracle 1e6", flush=True)
    E, cnt = _gpu_hist(geometry, source, material, O, 10_000_000, 51)
    Eg = E.cpu().numpy().view(np.uint32)
    assert int(Eg.sum(dtype=np.int64)) == int(cnt[_abi.CNT_DEPOSIT])
    med = osg.Medium(2, False, 5, [6, 3.46], [0.1, 0.1, 0.1], [0, 0, 0], 'gaussian', 0.5)
    hm = orc.host_material_from_medium(med, Q=500.0)
    Eo, _ = orc.forward(geometry, source, None, O, 1_000_000, seed=52, packet=10_000, host_material=hm)
    _compare_populated(Eg, 10_000_000, Eo, 1_000_000, "config 2 material")


def test_config3_material_against_oracle(gpu):
    geometry, source, material, observation = cases.config3(n=1)
    observation = dict(observation, time=np.arange(0, 3 + 1e-9, 0.05))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, False, observation, 1, allow_nd=True)
    print("config 3 material: GPU 1e7, oracle 4e5", flush=True)
    E, cnt = _gpu_hist(geometry, source, material, O, 10_000_000, 53)
    Eg = E.cpu().numpy().view(np.uint32)
    assert int(cnt[_abi.CNT_BOUNDARY]) > 0
    med = osg.Medium(3, False, 20, [6, 6 / math.sqrt(3)], [0.05, 0.05, 0.05], [0, 0, 0], 'VonKarman', 0.2, 0.5)
    hm = orc.host_material_from_medium(med, rho=2000.0)
    Eo, _ = orc.forward(geometry, source, None, O, 400_000, seed=54, packet=10_000, host_material=hm)
    _compare_populated(Eg, 10_000_000, Eo, 400_000, "config 3 material")


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_config5_material_against_oracle(gpu, precision):
    geometry, source, material, observation = cases.config5(n=1)
    observation = dict(observation, time=np.arange(0, 0.4 + 1e-9, 0.02))
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1, allow_nd=True)
    print(f"config 5 material ({precision}): GPU 2e7, oracle 3e5, {O.size} bins", flush=True)
    E, cnt = _gpu_hist(geometry, source, material, O, 20_000_000, 55, precision=precision)
    Eg = E.cpu().numpy().view(np.uint32)
    del E
    torch.cuda.empty_cache()
    med = osg.Medium(3, True, 40, 1, [0.1, 0.1], [0.0], 'exp', 0.05)
    hm = orc.host_material_from_medium(med, Q=300.0)
    Eo, _ = orc.forward(geometry, source, None, O, 300_000, seed=56, packet=10_000, host_material=hm)
    _compare_populated(Eg, 20_000_000, Eo, 300_000, f"config 5 material {precision}")


def test_config3_full_size_properties(gpu):
    """5e7 particles, the full 301 times, with z split at the free surface."""
    geometry, source, material, observation = cases.config3()
    observation = dict(observation, z=[-100.0, 0.0, 100.0])
    n = 50_000_000
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, False, observation, n, allow_nd=True)
    dm = solver.DeviceMaterial(material, 3, 'cuda', with_zoeppritz=True)
    print("config 3 full size: 5e7 particles, default placement then direct atomics", flush=True)
    E3, c3, _k = solver.forward_histogram(geometry, source, material, O, n, seed=57, device='cuda', dmat=dm)
    E1, c1, _k = solver.forward_histogram(geometry, source, material, O, n, seed=57, device='cuda', dmat=dm,
                                          hist_mode=1)
    torch.cuda.synchronize()
    assert torch.equal(c1, c3) and torch.equal(E1, E3)
    H = E3.cpu().numpy().view(np.uint32).reshape(O.shape, order='F')
    assert int(H.sum(dtype=np.int64)) == int(c3[_abi.CNT_DEPOSIT]) > 0
    assert int(c3[_abi.CNT_BOUNDARY]) > n // 10             # free-surface reflections
    # binned variables (r, z, direction), then time and mode: z >= 0 is never observed
    assert H.shape[1] == 2 and H[:, 1].sum() == 0 and H[:, 0].sum() > 0
    # Q = Inf and a half-space: no particle is lost, so each time slice holds at most n
    per_t = H.sum(axis=(0, 1, 2, 4), dtype=np.int64)
    assert np.all(per_t <= n) and per_t[1] == n


def test_config5_full_size_properties(gpu):
    """1e9 particles, 28.4 GB histogram: placement invariance, deposits ==
    counter, and for t <= 5.5 (no particle born within 0.3 of the origin can
    reach the grid's edge at 6.05) the count of each time slice is
    Binomial(N, exp(-omega t / Q)): survival of k sub-steps of probability
    exp(-omega dt_s / Q) each."""
    geometry, source, material, observation = cases.config5()
    n = 1_000_000_000
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, n, allow_nd=True)
    dm = solver.DeviceMaterial(material, 3, 'cuda')
    dbins = solver.DeviceBins(O, 'cuda')
    print("config 5 full size: 1e9 particles, time-major placement", flush=True)
    E3, c3, _k = solver.forward_histogram(geometry, source, material, O, n, seed=58, device='cuda', dmat=dm,
                                          bins=dbins)
    torch.cuda.synchronize()
    # E is column-major (x, y, z, dir, time): the time slice is a contiguous block
    ts = int(np.prod(O.shape[:4]))
    per_t = E3.view(O.Nt, ts).sum(dim=1, dtype=torch.int64).cpu().numpy()
    assert int(per_t.sum()) == int(c3[_abi.CNT_DEPOSIT])
    omega = 2 * math.pi * 40.0
    checked = 0
    for it in range(1, O.Nt):
        t = O.t[it]
        if t > 5.5:
            break
        p = math.exp(-omega * t / 300.0)
        sd = math.sqrt(n * p * (1 - p))
        assert abs(per_t[it] - n * p) < 5 * sd + 1, (t, per_t[it], n * p)
        checked += 1
    assert checked > 250
    print("config 5 full size: direct atomics", flush=True)
    E1, c1, _k = solver.forward_histogram(geometry, source, material, O, n, seed=58, device='cuda', dmat=dm,
                                          bins=dbins, hist_mode=1)
    torch.cuda.synchronize()
    assert torch.equal(c1, c3)
    assert torch.equal(E1, E3)

"""Elastic boundary partitions on the GPU (the cases of
Examples/BoundaryReflectionEnergyTests.m, tests/boundary_cases.py).

- Function-level kernel (rtmc_propagate_smalldt): the same pure P/SV/SH
  packets as the oracle pins, plane z = 0 and cylinder; energy fractions
  within 4 binomial sigma of the code's expectation, mirror and Snell
  directions.
- Tracing kernel (rtmc_forward): plane-wave sources at the reference's
  angles onto the free plane z = 0, no volume scattering.  A P source ends
  with E_P = E_Rpp(theta); an S source, whose polarisation is uniform in the
  wavefront (initializeParticle.m:146-147), is SV with probability 1/2, so it
  ends with E_P = (1 - E_Rsvsv(theta)) / 2 below the critical angle and 0
  beyond it.  A point source on the axis of a cylinder: each ray from the
  axis meets the wall at incidence acos(sin phi) (phi from the axis) when
  v t sin(phi) >= R, once before t = 3R / v, so
  E_P(t) = 1 - int_{sin phi >= R/(v t)} (1 - E_Rpp(acos(sin phi))) sin(phi) / 2 dphi.
"""
import math

import numpy as np
import pytest
import torch

from radiativetransfermontecarlo_amd import _abi, solver
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import boundary_cases as bc

pytestmark = pytest.mark.gpu

N = 200_000
DT = 0.01


class _DevMat:
    def __init__(self, dev):
        self.arrays = {k: torch.tensor(v, dtype=torch.float64, device=dev) for k, v in bc.material_arrays().items()}
        self.c = bc.material_struct(self.arrays, lambda t: t.data_ptr())


def _smalldt(packet, geometry, dev):
    x, d, perp, p = packet
    f = dict(dtype=torch.float64, device=dev)
    n = p.size
    P = solver.Particles(3, n, torch.tensor(x, **f), torch.tensor(d, **f), torch.tensor(perp, **f),
                         torch.tensor(p, dtype=torch.uint8, device=dev), torch.ones(n, dtype=torch.uint8, device=dev),
                         torch.zeros(n, **f))
    info = solver.propagateParticleSmallDt(_DevMat(dev), geometry, P, DT, solver.Draws.philox(3, 1), T_eff=DT)
    torch.cuda.synchronize()
    assert info.n_sub == 1
    return {k: getattr(P, k).cpu().numpy() for k in ('x', 'dir', 'perp', 'p')}


PLANE = {'dimension': 3, 'frame': 'cartesian', 'bnd': [{'dir': 3, 'val': 0.0}]}
CYL = {'dimension': 3, 'frame': 'cartesian', 'bnd': [{'dir': 4, 'val': 1000.0}]}
CASES = [('P', a) for a in bc.P_ANGLES] + [('SV', a) for a in bc.SV_ANGLES] + [('SH', a) for a in bc.SH_ANGLES]


@pytest.mark.parametrize("mode,deg", CASES)
def test_function_kernel_plane_partition(gpu, mode, deg):
    pk = bc.plane_packet(mode, deg, N, np.random.default_rng(int(deg) + len(mode)), DT)
    G = _smalldt(pk, PLANE, gpu)
    n = np.repeat(np.array([[0.0], [0.0], [1.0]]), N, axis=1)
    assert np.all(G['x'][2] > 0)
    bc.check_partition(mode, deg, pk[0], pk[1], G['dir'], G['perp'], G['p'], N, n, bc.zoeppritz())


@pytest.mark.parametrize("mode,deg", [('P', 15.0), ('P', 60.0), ('SV', 20.0), ('SV', 30.0), ('SH', 45.0)])
def test_function_kernel_cylinder_partition(gpu, mode, deg):
    pk = bc.cylinder_packet(mode, deg, N, np.random.default_rng(100 + int(deg)), DT)
    G = _smalldt(pk, CYL, gpu)
    r = np.hypot(G['x'][0], G['x'][1])
    assert np.all(r < 1000.0)
    n = np.stack([G['x'][0] / r, G['x'][1] / r, np.zeros(N)])
    bc.check_partition(mode, deg, pk[0], pk[1], G['dir'], G['perp'], G['p'], N, n, bc.zoeppritz())


def _trace_modes(geometry, source, observation, n, dev, seed):
    """Fraction of P particles at every observation time (one bin per variable)."""
    O = initializeObservation(geometry, False, observation, n, allow_nd=True)
    dm = _DevMat(dev)
    E, cnt, _k = solver.forward_histogram(geometry, source, None, O, n, seed=seed, device=dev, dmat=dm)
    H = E.cpu().numpy().view(np.uint32).reshape(O.shape, order='F')
    per = H.reshape(-1, O.Nt, 2).sum(axis=0)        # (time, mode)
    return O, per[:, 0] / n, per.sum(axis=1), cnt.cpu().numpy()


@pytest.mark.parametrize("pol,deg", [('P', a) for a in bc.P_ANGLES] + [('S', a) for a in bc.SV_ANGLES])
def test_tracing_kernel_plane_source_partition(gpu, pol, deg):
    th = math.radians(deg)
    n = 2_000_000
    v = bc.VP if pol == 'P' else bc.VS
    z0 = 80.0
    src = {'type': 'plane', 'direction': [math.sin(th), 0.0, -math.cos(th)], 'position': [0.0, 0.0, z0],
           'extent': [20.0, 20.0], 'lambda': 1e-3, 'polarization': pol}
    t_end = 1.25 * (z0 + 15.0) / (v * math.cos(th))
    big = [-1e4, 1e4]
    observation = {'x': big, 'y': big, 'z': big, 'directions': [0, math.pi], 'time': np.linspace(0, t_end, 11)}
    O, EP, tot, cnt = _trace_modes(PLANE, src, observation, n, gpu, 61 + int(deg))
    assert np.all(tot[1:] == n)
    tab = bc.zoeppritz()
    if pol == 'P':
        want = bc.interp(tab, 'E_Rpp', deg)
    else:
        want = 0.0 if deg >= bc.CRIT else (1 - bc.interp(tab, 'E_Rsvsv', deg)) / 2
    sd = math.sqrt(max(want * (1 - want), 1.0 / n) / n)
    assert abs(EP[-1] - want) < 4 * sd + 1e-12, (pol, deg, EP[-1], want)
    assert int(cnt[_abi.CNT_BOUNDARY]) == n


def test_tracing_kernel_cylinder_partition(gpu):
    R = 1000.0
    geometry = dict(CYL, bnd=[{'dir': 4, 'val': R}])
    n = 4_000_000
    times = np.linspace(0, 2.9 * R / bc.VP, 30)
    big = [-1e5, 1e5]
    observation = {'x': big, 'y': big, 'z': big, 'directions': [0, math.pi], 'time': times}
    O, EP, tot, cnt = _trace_modes(geometry, {'lambda': 1e-6 * R}, observation, n, gpu, 71)
    assert np.all(tot[1:] == n)
    tab = bc.zoeppritz()
    phi = np.linspace(0, math.pi, 400_001)
    w = np.sin(phi) / 2 * (phi[1] - phi[0])
    inc = np.degrees(np.arccos(np.clip(np.sin(phi), -1, 1)))
    loss = 1 - np.interp(inc, tab['j1_deg'], tab['E_Rpp'])
    for it in (12, 20, 29):
        t = O.t[it]
        hit = np.sin(phi) * bc.VP * t > R
        want = 1 - float(np.sum((loss * w)[hit]))
        sd = math.sqrt(want * (1 - want) / n)
        assert abs(EP[it] - want) < 4 * sd + 1e-5, (t, EP[it], want)

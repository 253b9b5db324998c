"""The LDS budget of the forward plan at its edge.

The tracing kernel declares about 0.2 KB of LDS statically beside the block the
plan sizes (bin edges + histogram).  A histogram whose edges and counters fit
the work-group's LDS but not beside that static part must not be placed in
LDS: forced there (hist_mode 2) the call fails with an error, and by default
it goes to global memory with the same counts as direct global atomics.  At
the largest size that does fit, the LDS placement gives those counts too.
"""
import numpy as np
import pytest
import torch

from radiativetransfermontecarlo_amd import _abi, solver
from radiativetransfermontecarlo_amd.material import MaterialClass
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import cases

pytestmark = pytest.mark.gpu

STATIC_RESERVE = 256   # kTraceStaticLds: what the plan keeps back for the kernel's own LDS
N_TIMES = 10


def _case(nx):
    geometry = {'dimension': 3, 'frame': 'cartesian'}
    source = {'numberParticles': 1, 'type': 'point', 'position': [0, 0, 0], 'lambda': 1e-3}
    big = [-1e6, 1e6]
    observation = {'x': np.linspace(-6.0, 6.0, nx), 'y': big, 'z': big, 'directions': [0, np.pi],
                   'time': np.linspace(0, 2.0, N_TIMES)}
    material = MaterialClass.prepareSigma(cases.iso_acoustic(), 3)
    O = initializeObservation(geometry, True, observation, 1, allow_nd=True)
    lds = 8 * sum(int(O.edges[k].size) for k in range(4)) + 4 * O.size
    return geometry, source, material, O, lds


def _nx_for(target):
    """Largest x-edge count whose edges + histogram take at most `target` bytes."""
    lo, hi = 2, 8192
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if _case(mid)[4] <= target:
            lo = mid
        else:
            hi = mid
    return lo


def _counts(case, hist_mode, n=200_000, seed=5):
    geometry, source, material, O, _ = case
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, n, seed=seed, device='cuda',
                                          hist_mode=hist_mode)
    torch.cuda.synchronize()
    return E.cpu().numpy(), cnt.cpu().numpy()


def _max_lds():
    return int(torch.cuda.get_device_properties(0).shared_memory_per_block)


def test_histogram_beside_static_lds_goes_to_global(gpu):
    limit = _max_lds()
    case = _case(_nx_for(limit - 16))
    lds = case[4]
    assert limit - STATIC_RESERVE < lds <= limit, (lds, limit)
    print(f"edges + histogram {lds} B of {limit} B LDS", flush=True)
    with pytest.raises(_abi.RTMCError, match="LDS histogram"):
        _counts(case, 2)
    E0, c0 = _counts(case, 0)
    E1, c1 = _counts(case, 1)
    assert np.array_equal(E0, E1) and np.array_equal(c0, c1)
    assert int(E0.astype(np.int64).sum()) > 0


def test_largest_lds_histogram(gpu):
    limit = _max_lds()
    case = _case(_nx_for(limit - STATIC_RESERVE))
    assert case[4] <= limit - STATIC_RESERVE
    print(f"edges + histogram {case[4]} B of {limit} B LDS", flush=True)
    E2, c2 = _counts(case, 2)
    E1, c1 = _counts(case, 1)
    assert np.array_equal(E2, E1) and np.array_equal(c2, c1)
    assert int(E2.astype(np.int64).sum()) > 0

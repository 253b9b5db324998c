"""Edge cases of the C ABI (rtmc_forward, rtmc_forward_batched, rtmc_adjoint_batched):
empty and ragged inputs, a single observation time, bins no particle reaches,
and the host-side argument checks (they return an error status and message
before anything is launched)."""
import ctypes as C
import math
import re

import numpy as np
import pytest
import torch

from radiativetransfermontecarlo_amd import _abi
from radiativetransfermontecarlo_amd import batch as rb
from radiativetransfermontecarlo_amd import solver
from radiativetransfermontecarlo_amd.material import MaterialClass
from radiativetransfermontecarlo_amd.observation import initializeObservation
from tests import cases

pytestmark = pytest.mark.gpu


def _setup(time=np.arange(0, 0.5 + 1e-9, 0.02), x=np.arange(0, 6 + 1e-9, 0.05)):
    geometry, source, material, observation = cases.config1()
    observation = dict(observation, time=time, x=x)
    material = MaterialClass.prepareSigma(material, 3)
    O = initializeObservation(geometry, True, observation, 1)
    return geometry, source, material, O


def test_zero_particles_leaves_histogram_untouched(gpu):
    geometry, source, material, O = _setup()
    E = torch.full((O.size,), 7, dtype=torch.int32, device='cuda')
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, 0, seed=1, device='cuda', E=E)
    assert torch.all(E == 7)
    assert int(cnt.abs().sum()) == 0


def test_single_observation_time_has_no_deposits(gpu):
    """The histogram slice of t = 0 is never filled (the first deposit is at the
    second observation time); with one time there is nothing to deposit."""
    geometry, source, material, O = _setup(time=np.array([0.0]))
    assert O.Nt == 1
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, 10_000, seed=2, device='cuda')
    c = cnt.cpu().numpy()
    assert int(E.sum()) == 0 and c[_abi.CNT_DEPOSIT] == 0


def test_bins_out_of_reach_count_nothing(gpu):
    """Radii 10..20 cannot be reached by t = 0.5 at speed 2: no deposit, but the
    particles still scatter."""
    geometry, source, material, O = _setup(x=np.linspace(10, 20, 11))
    E, cnt, _k = solver.forward_histogram(geometry, source, material, O, 20_000, seed=3, device='cuda')
    c = cnt.cpu().numpy()
    assert int(E.sum()) == 0 and c[_abi.CNT_DEPOSIT] == 0
    assert c[_abi.CNT_SCATTER] > 0


def test_ragged_batch_with_empty_items(gpu):
    geometry, source, base, observation = cases.config1()
    observation = dict(observation, time=np.arange(0, 0.6 + 1e-9, 0.02))
    params = [(0.05, 0.1, math.inf), (0.1, 0.3, 200.0), (0.05, 0.2, math.inf), (0.08, 0.15, 500.0)]
    mats = [rb.prepare_material(base, 3, c=c, lc=lc, Q=Q) for c, lc, Q in params]
    O = initializeObservation(geometry, True, observation, 1)
    Ns, seeds = [0, 31_337, 0, 1], [1, 2, 3, 4]
    B = rb.Batch(geometry, O, mats, source, Ns, seeds, device='cuda')
    E, cnt = B.forward()
    for i, m in enumerate(mats):
        if Ns[i] == 0:
            assert int(E[i].abs().sum()) == 0 and int(cnt[i].abs().sum()) == 0, f"item {i}"
            continue
        Es, cs, _k = solver.forward_histogram(geometry, source, m, O, Ns[i], seed=seeds[i], device='cuda')
        assert torch.equal(E[i], Es) and torch.equal(cnt[i], cs), f"item {i}"
    # the adjoint of the same ragged batch: empty items get a zero gradient
    w = torch.ones((len(mats), O.size), dtype=torch.float32, device='cuda')
    g = B.adjoint(w)
    g = np.asarray(g.cpu() if hasattr(g, 'cpu') else g)
    assert np.all(g[0] == 0.0) and np.all(g[2] == 0.0)
    assert np.all(np.isfinite(g))


def test_empty_batch_is_a_no_op(gpu):
    lib = _abi.load()
    geometry, source, material, O = _setup()
    b = _abi.Batch()
    b.geometry = solver.build_geometry(geometry)
    bins = solver.DeviceBins(O, 'cuda')
    b.bins = bins.c
    b.n_times = O.Nt
    b.dt = 0.02
    b.n_items = 0
    b.items = None
    st = torch.cuda.current_stream().cuda_stream
    assert lib.rtmc_forward_batched(C.byref(b), C.c_void_p(st)) == _abi.OK


def _status(fn, *args):
    lib = _abi.load()
    rc = fn(*args)
    return rc, lib.rtmc_last_error().decode(errors="replace")


@pytest.mark.parametrize("field,value,msg", [
    ("n_particles", -5, "n_particles must be >= 0"),
    ("hist_mode", 7, "hist_mode must be"),
    ("n_times", 0, "n_times must be >= 1"),
    ("block_size", 100, "block_size must be a multiple of 64"),
    ("precision", 9, "precision"),
])
def test_invalid_arguments_are_rejected_before_launch(gpu, field, value, msg):
    lib = _abi.load()
    geometry, source, material, O = _setup()
    keep = []
    prob = _abi.Problem()
    prob.geometry = solver.build_geometry(geometry)
    prob.source = solver.build_source(source, 3, keep, torch.device('cuda'))
    dmat = solver.DeviceMaterial(material, 3, 'cuda')
    prob.material = dmat.c
    bins = solver.DeviceBins(O, 'cuda')
    prob.bins = bins.c
    prob.n_particles = 1000
    prob.seed = 1
    prob.n_times = O.Nt
    prob.dt = 0.02
    setattr(prob, field, value)
    E = torch.zeros(O.size, dtype=torch.int32, device='cuda')
    cnt = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device='cuda')
    st = torch.cuda.current_stream().cuda_stream
    rc, err = _status(lib.rtmc_forward, C.byref(prob), C.c_void_p(E.data_ptr()), C.c_void_p(cnt.data_ptr()),
                      C.c_void_p(st))
    torch.cuda.synchronize()
    assert rc != _abi.OK and msg in err, (rc, err)
    assert int(E.abs().sum()) == 0


def test_null_pointers_are_rejected(gpu):
    lib = _abi.load()
    st = torch.cuda.current_stream().cuda_stream
    rc, err = _status(lib.rtmc_forward, None, None, None, C.c_void_p(st))
    assert rc != _abi.OK and "must be given" in err
    rc, err = _status(lib.rtmc_forward_batched, None, C.c_void_p(st))
    assert rc != _abi.OK and "batch must be given" in err


@pytest.mark.parametrize("var,value,msg", [
    ("RTMC_HOT_SLOTS", "1", "RTMC_HOT_SLOTS must be 0 or a power of two in [2, 8192]"),
    ("RTMC_HOT_SLOTS", "3000", "RTMC_HOT_SLOTS must be 0 or a power of two in [2, 8192]"),
    ("RTMC_RUN_PARTICLES", "0", "RTMC_RUN_PARTICLES must be in [1, 2097151]"),
    ("RTMC_RUN_PARTICLES", "2097152", "RTMC_RUN_PARTICLES must be in [1, 2097151]"),
    ("RTMC_GUIDE_FINER", "7", "RTMC_GUIDE_FINER must be -1 (automatic) or in [0, 6]"),
])
def test_invalid_tuning_variables_are_rejected_before_launch(gpu, monkeypatch, var, value, msg):
    """The time-major tuning variables are checked when the launch is planned:
    an error status and message, and E untouched."""
    geometry = {'dimension': 2, 'frame': 'cartesian'}
    source = {'numberParticles': 1, 'lambda': 0.01, 'position': [0, 0]}
    e = np.linspace(-2, 2, 9)
    observation = {'x': e, 'y': e, 'time': np.arange(0, 0.3 + 1e-9, 0.05)}
    m = MaterialClass.prepareSigma(cases.elastic_iso(2), 2)
    O = initializeObservation(geometry, False, observation, 1)
    monkeypatch.setenv(var, value)
    E = torch.zeros(O.size, dtype=torch.int32, device='cuda')
    with pytest.raises(RuntimeError, match=re.escape(msg)):
        solver.forward_histogram(geometry, source, m, O, 1000, seed=1, device='cuda', hist_mode=3, E=E)
    torch.cuda.synchronize()
    assert int(E.abs().sum()) == 0

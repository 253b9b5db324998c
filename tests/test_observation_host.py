"""Host logic: initializeObservation and the normalisation (CPU only), plus the
C ABI library loading and exporting every symbol of include/rtmc.h."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from radiativetransfermontecarlo_amd import _abi
from radiativetransfermontecarlo_amd.observation import initializeObservation, normalize

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_spherical_volume_weights_closed_form():
    """3-D spherical bin weight r^2 dr + dr^3/12 = ((r+dr/2)^3 - (r-dr/2)^3)/3."""
    O = initializeObservation({'dimension': 3, 'frame': 'spherical'}, True,
                              {'x': np.arange(0, 6.0001, 0.05), 'time': np.arange(0, 1.01, 0.02)}, 100)
    e = O.obs['binX']
    np.testing.assert_allclose(O.obs['dx'], (e[1:] ** 3 - e[:-1] ** 3) / 3, rtol=1e-12)
    assert O.ibins == [1, 2]
    assert O.shape == (120, 1, 51, 1)
    assert O.obs['binZ'][0] == pytest.approx(-1) and O.obs['binZ'][1] == pytest.approx(1)


def test_cylindrical_and_2d_weights():
    O = initializeObservation({'dimension': 3, 'frame': 'cylindrical'}, False,
                              {'x': [0, 1, 2], 'time': [0, 1]}, 10)
    np.testing.assert_allclose(O.obs['dx'], [0.5 * 1, 1.5 * 1])       # r dr (e^2 - e^2)/2
    O2 = initializeObservation({'dimension': 2, 'frame': 'spherical'}, True, {'x': [0, 1, 3], 'time': [0, 1]}, 10)
    np.testing.assert_allclose(O2.obs['dx'], [0.5, 2 * 2])
    assert O2.n_modes == 1
    assert O.n_modes == 2 and O.shape[-1] == 2


def test_time_vector_and_direction_bins():
    O = initializeObservation({'dimension': 3, 'frame': 'spherical'}, True,
                              {'x': [0, 1], 'directions': [0, np.pi / 2, np.pi], 'time': [0.5, 0.1, 0.3]}, 10)
    np.testing.assert_allclose(O.t, [0, 0.1, 0.3, 0.5])
    np.testing.assert_allclose(O.obs['binPsi'], [-1, 0, 1], atol=1e-15)   # cos of reversed directions
    assert O.ibins == [1, 4]


def test_more_than_two_binned_variables():
    obs = {'x': [-1, 0, 1], 'y': [-1, 0, 1], 'directions': np.linspace(0, np.pi, 4), 'time': [0, 1]}
    with pytest.raises(ValueError, match='two directions'):
        initializeObservation({'dimension': 2, 'frame': 'cartesian'}, True, obs, 10)
    O = initializeObservation({'dimension': 2, 'frame': 'cartesian'}, True, obs, 10, allow_nd=True)
    assert O.ibins == [1, 2, 4]
    assert O.strides == [1, 2, 0, 4]
    assert O.shape == (2, 2, 3, 2, 1)


def test_normalisation_rule():
    """energyDensity = E/(N dx dy dz); the direction width cancels (radiativeTransfer.m:161-188)."""
    geo = {'dimension': 3, 'frame': 'spherical'}
    obs = {'x': [0, 1, 2], 'directions': [0, np.pi / 3, np.pi], 'time': [0, 1]}
    O = initializeObservation(geo, True, obs, 4)
    assert O.ibins == [1, 4]
    E = np.arange(O.size, dtype=np.uint32)
    dens, energy = normalize(O, E, 4)
    Ef = E.reshape(O.shape, order='F')[..., 0].astype(float)
    dx = O.obs['dx']
    dy = O.obs['dy'][0]
    dz = O.obs['dz'][0]
    dpsi = O.obs['dpsi']
    # the reference formula for ibins [1 4]: E ./ (d1' * (d2 * N)) .* dpsi / (dy * dz)
    ref = Ef / (dx[:, None, None] * (dpsi[None, :, None] * 4)) * dpsi[None, :, None] / (dy * dz)
    np.testing.assert_allclose(dens, ref, rtol=1e-14)
    np.testing.assert_allclose(energy, Ef.sum(axis=(0, 1)) / 4)


def test_library_exports_every_header_symbol():
    lib_path = _abi.LIB_PATH
    if not os.path.exists(lib_path):
        pytest.skip("librtmc.so not built")
    hdr = open(os.path.join(ROOT, 'include', 'rtmc.h')).read()
    declared = set(re.findall(r'^\s*(?:const\s+)?\w+\s*\*?\s*(rtmc_\w+)\s*\(', hdr, flags=re.M))
    assert declared == set(_abi.EXPORTS)
    lib = ctypes.CDLL(lib_path)
    for name in declared:
        assert hasattr(lib, name), name
    lib = _abi.load()
    assert lib.rtmc_abi_version() == _abi.ABI_VERSION


def test_struct_layouts_match_header():
    """ctypes mirrors have the sizes the C compiler gives the header structs."""
    import subprocess
    import tempfile
    src = r'''
#include <stdio.h>
#include "rtmc.h"
int main(){printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(rtmc_geometry),
 sizeof(rtmc_table), sizeof(rtmc_material), sizeof(rtmc_source), sizeof(rtmc_particles), sizeof(rtmc_draws),
 sizeof(rtmc_bins), sizeof(rtmc_problem), sizeof(rtmc_substep_info), sizeof(rtmc_batch_item), sizeof(rtmc_batch),
 sizeof(rtmc_adjoint_item), sizeof(rtmc_tuning), sizeof(rtmc_material_params), sizeof(rtmc_material_host),
 sizeof(rtmc_lds_knot), sizeof(rtmc_lds_table));
 return 0;}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, 'l.c')
        open(c, 'w').write(src)
        exe = os.path.join(d, 'l')
        subprocess.run(['gcc', '-I', os.path.join(ROOT, 'include'), c, '-o', exe], check=True)
        sizes = [int(v) for v in subprocess.run([exe], capture_output=True, text=True).stdout.split()]
    mine = [ctypes.sizeof(t) for t in (_abi.Geometry, _abi.Table, _abi.Material, _abi.Source, _abi.Particles,
                                       _abi.Draws, _abi.Bins, _abi.Problem, _abi.SubstepInfo, _abi.BatchItem,
                                       _abi.Batch, _abi.AdjointItem, _abi.Tuning, _abi.MaterialParams,
                                       _abi.MaterialHost, _abi.LdsKnot, _abi.LdsTable)]
    assert sizes == mine

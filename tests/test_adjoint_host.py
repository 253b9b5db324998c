"""Host-side derivative data of the adjoint (radiativetransfermontecarlo_amd.batch).

The likelihood-ratio score of a scattering angle is
dlogsigma_dlc(theta) - dlogSigma_dlc.  Pinned here against the table the
kernels sample from: d/dlc of E_T[cos theta] (FD of the table rebuilt at
lc +- h) must equal E_T[cos theta * score], and E_T[score] = 0; and against
the closed form of the exponential ACF in 3-D (MaterialClass.m:204-235,
:479-710): sigma ~ zeta^3 / (1 + (zeta s / a)^2)^2, s = 2 sin(theta/2).
"""
import math

import numpy as np

from radiativetransfermontecarlo_amd import batch as rb
from tests import cases


def _base():
    geometry, source, base, observation = cases.config1()
    return base


def test_dlogsigma_closed_form_exp_3d():
    m = rb.prepare_material(_base(), 3, c=0.08, lc=0.2, Q=math.inf)
    g, dS = rb.lc_derivatives(m, 3)
    th = np.linspace(0, np.pi, g.size)
    z = 2 * np.pi * 10 * 0.2 * np.sqrt(2 * (1 - np.cos(th))) / 6 ** (1 / 3)
    ga = 3 / 0.2 - 4 * z ** 2 / (1 + z ** 2) / 0.2
    assert np.max(np.abs(g - ga)) < 1e-6


def test_angle_score_matches_table_derivative():
    base = _base()
    c, lc, M = 0.08, 0.2, 1_000_000
    u = (np.arange(M) + 0.5) / M

    def table(lcv):
        return rb.prepare_material(base, 3, c=c, lc=lcv, Q=math.inf, knot_threshold=0.0)

    m = table(lc)
    g, dS = rb.lc_derivatives(m, 3)
    th = m.invcdf(u)
    score = np.interp(th, np.linspace(0, np.pi, g.size), g) - dS
    assert abs(np.mean(score)) < 1e-3 * np.std(score)
    h = 1e-3 * lc
    fd = (np.cos(table(lc + h).invcdf(u)).mean() - np.cos(table(lc - h).invcdf(u)).mean()) / (2 * h)
    lr = np.mean(np.cos(th) * score)
    assert abs(lr - fd) < 2e-3 * abs(fd)


def test_full_tables_follow_sigma():
    """knot_threshold=0 keeps a knot wherever the cdf grows; the reference's
    1e-8 rule stops at the last knot above it and extrapolates."""
    base = _base()
    full = rb.prepare_material(base, 3, c=0.08, lc=0.2, Q=math.inf, knot_threshold=0.0)
    ref = rb.prepare_material(base, 3, c=0.08, lc=0.2, Q=math.inf)
    assert np.all(np.diff(full.invcdf.x) > 0)
    assert full.invcdf.y[-1] > 3.14 and full.invcdf.x[-1] > 1 - 1e-5
    assert full.invcdf(np.array([1.0]))[0] == full.invcdf.y[-1]       # held, not extrapolated
    assert ref.invcdf.y[-1] < 2.0 and ref.invcdf.x[-1] < 0.999        # truncated tail
    assert full.Sigma == ref.Sigma

"""Compressed inverse-CDF tables (rtmc_compress_table), the form in which the
tracing kernel holds the scattering-angle tables in LDS.

The bound rtmc.h states, |theta(q) - table(q)| <= tol + 4e-7 for every q, is
checked by brute force against numpy's own linear interpolation of the full
table (np.interp, with the table's linear or nearest extrapolation written
out), at every knot of the full table, at every midpoint between two of its
knots, at 2e5 random q and at the ends of (0, 1].  The compressed table is
evaluated the way the kernel evaluates it: the guide over 2^g cells brackets
the knot, a binary search inside the bracket finds it, theta and slope are
float32 and the chord is computed in float64.  The tables are the reference's
own (MaterialClass.prepareSigmaOne, 1e6-point angle grid) for the workloads
the benchmarks run."""
import math

import numpy as np
import pytest

from radiativetransfermontecarlo_amd import native_material as nm
from radiativetransfermontecarlo_amd.material import MaterialClass
from tests import cases

FLOAT_ALLOWANCE = 4e-7


def _tables(name):
    """(x, y, extrap) of every channel of a workload's material."""
    if name == "readme":
        g, _s, m, _o = cases.config1()
    elif name == "large3d":
        g, _s, m, _o = cases.config5()
    elif name == "elastic2d":
        g, _s, m, _o = cases.config2()
    elif name == "halfspace":
        g, _s, m, _o = cases.config3()
    elif name == "readme_nearest":
        g, _s, m, _o = cases.config1()
        p = nm.params_from(m, 3, knot_threshold=0.0)
        h = nm.prepare_host(p)
        return [(x, y, h["extrap"]) for x, y in h["tables"]]
    p = nm.params_from(m, g['dimension'])
    h = nm.prepare_host(p)
    return [(x, y, h["extrap"]) for x, y in h["tables"]]


def _full(x, y, extrap, q):
    """griddedInterpolant(x, y) at q: np.interp inside, the end intervals' lines
    (linear) or the end values (nearest) outside."""
    v = np.interp(q, x, y)
    if extrap == "linear":
        lo = q < x[0]
        hi = q > x[-1]
        v[lo] = y[0] + (q[lo] - x[0]) * (y[1] - y[0]) / (x[1] - x[0])
        v[hi] = y[-2] + (q[hi] - x[-2]) * (y[-1] - y[-2]) / (x[-1] - x[-2])
    return v


def _kernel_eval(kn, extrap, q, glog2):
    """The kernel's lookup (rtmc_trace.hpp): clamp (nearest), guide cell,
    bracketed binary search, float32 theta + (q - u) * float32 slope."""
    u = kn["u"]
    m = u.size
    G = 1 << glog2
    # guide: cell c holds the largest k <= m-2 with u_k <= c / G (0 below u_0)
    cq = np.arange(G + 1) / G
    guide = np.clip(np.searchsorted(u[:m - 1], cq, side="right") - 1, 0, m - 2)
    if extrap == "nearest":
        q = np.clip(q, u[0], u[-1])
    c = np.clip((q * G).astype(np.int64), 0, G - 1)
    lo = guide[c].copy()
    hi = guide[c + 1] + 1
    while True:
        act = hi - lo > 1
        if not act.any():
            break
        mid = (lo + hi) >> 1
        go = act & (u[np.minimum(mid, m - 1)] <= q)
        lo = np.where(go, mid, lo)
        hi = np.where(act & ~go, mid, hi)
    # the bracketed search must find what a plain search over all knots finds
    k_ref = np.clip(np.searchsorted(u, q, side="right") - 1, 0, m - 2)
    assert np.array_equal(lo, k_ref)
    th = kn["theta"].astype(np.float64)[lo]
    sl = kn["slope"].astype(np.float64)[lo]
    return th + (q - u[lo]) * sl


def _queries(x, rng):
    mids = 0.5 * (x[1:] + x[:-1])
    ends = np.array([2.0 ** -32, 1e-12, x[0] * 0.5, x[-1] + 0.5 * (1 - x[-1]), 1.0])
    return np.concatenate([x, mids, rng.random(200_000), ends])


@pytest.mark.parametrize("name", ["readme", "large3d", "elastic2d", "halfspace", "readme_nearest"])
@pytest.mark.parametrize("tol", [1e-5, 1e-4])
def test_compressed_table_within_tol_everywhere(name, tol):
    rng = np.random.default_rng(7)
    for x, y, extrap in _tables(name):
        kn, used = nm.compress_table(x, y, tol, 65535)
        assert used == tol
        u = kn["u"]
        m = u.size
        # a subset of the table's own knots, the first two and the last two included
        assert m >= 4 and np.all(np.diff(u) > 0)
        idx = np.searchsorted(x, u)
        assert np.array_equal(x[idx], u)
        assert idx[0] == 0 and idx[1] == 1 and idx[-2] == x.size - 2 and idx[-1] == x.size - 1
        assert np.array_equal(kn["theta"], y[idx].astype(np.float32))
        q = _queries(x, rng)
        glog2 = max(6, min(12, int(math.ceil(math.log2(m)))))
        err = np.abs(_kernel_eval(kn, extrap, q, glog2) - _full(x, y, extrap, q))
        assert err.max() <= tol + FLOAT_ALLOWANCE, (name, err.max(), m)
        # far fewer knots than the table (it has 3e5 - 1e6)
        assert m < x.size / 200


def test_compression_doubles_tol_to_fit_max_knots():
    x, y, extrap = _tables("large3d")[0]
    kn, used = nm.compress_table(x, y, 1e-6, 64)
    assert kn.size <= 64 and used > 1e-6
    kn2, used2 = nm.compress_table(x, y, used / 2, 65535)
    assert kn2.size > 64  # half the achieved tolerance would not have fitted
    q = _queries(x, np.random.default_rng(3))
    err = np.abs(_kernel_eval(kn, extrap, q, 6) - _full(x, y, extrap, q))
    assert err.max() <= used + FLOAT_ALLOWANCE


def test_compression_of_a_straight_table_keeps_the_four_end_knots():
    """Collinear knots (the isotropic 2-D law theta = pi u) all drop out (tol
    1e-12 covers the rounding of linspace)."""
    x = np.linspace(0.0, 1.0, 10_001)
    y = np.pi * x
    kn, used = nm.compress_table(x, y, 1e-12, 1024)
    assert kn.size == 4 and used == 1e-12
    np.testing.assert_array_equal(kn["u"], x[[0, 1, -2, -1]])


def test_isotropic_3d_table_against_acos():
    """theta = acos(1 - 2u) (the isotropic 3-D law): the compressed table of a
    1e5-knot table stays within tol of the closed form up to the table's own
    interpolation error."""
    x = np.linspace(0.0, 1.0, 100_001)
    y = np.arccos(1.0 - 2.0 * x)
    tol = 1e-4
    kn, used = nm.compress_table(x, y, tol, 65535)
    q = np.random.default_rng(5).random(100_000) * (1 - 2e-3) + 1e-3
    own = np.abs(np.interp(q, x, y) - np.arccos(1.0 - 2.0 * q)).max()
    err = np.abs(_kernel_eval(kn, "linear", q, 10) - np.arccos(1.0 - 2.0 * q)).max()
    assert err <= tol + own + FLOAT_ALLOWANCE


@pytest.mark.parametrize("bad", ["decreasing", "nan", "few", "tol"])
def test_compression_rejects_bad_tables(bad):
    from radiativetransfermontecarlo_amd import _abi
    x = np.linspace(0, 1, 100)
    y = np.sqrt(x)
    tol, mk = 1e-5, 1024
    if bad == "decreasing":
        x = x[::-1].copy()
    elif bad == "nan":
        y = y.copy()
        y[50] = np.nan
    elif bad == "few":
        mk = 3
    else:
        tol = -1.0
    with pytest.raises(_abi.RTMCError):
        nm.compress_table(x, y, tol, mk)


def test_attach_rejects_materials_without_tables():
    """rtmc_attach_lds_tables checks its materials before touching the device
    (no GPU needed): a material without an inverse-CDF table is an argument
    error, and nothing is attached."""
    import ctypes as C
    from radiativetransfermontecarlo_amd import _abi
    lib = _abi.load()
    mats = (_abi.Material * 2)()
    mats[0].acoustics = 1
    mats[1].acoustics = 0
    h = C.c_void_p()
    rc = lib.rtmc_attach_lds_tables(mats, 2, 1e-5, 0, C.byref(h), None)
    assert rc == _abi.ERR_ARG and not h.value
    assert "inverse-CDF table 0 is missing" in lib.rtmc_last_error().decode()
    assert all(mats[i].lds[c].m == 0 for i in range(2) for c in range(4))
    assert lib.rtmc_attach_lds_tables(mats, 0, 1e-5, 0, C.byref(h), None) == _abi.OK and not h.value
    assert lib.rtmc_attach_lds_tables(mats, 1, -1.0, 0, C.byref(h), None) == _abi.ERR_ARG
    assert lib.rtmc_attach_lds_tables(mats, 1, 1e-5, 3, C.byref(h), None) == _abi.ERR_ARG

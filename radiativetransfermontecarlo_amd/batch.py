"""Batched forward, adjoint and autotuner (rtmc_forward_batched,
rtmc_adjoint_batched, rtmc_autotune_batched).

The reference solves one material per ``radiativeTransfer`` call
(radiativeTransfer.m:1-207) and has no derivatives.  A :class:`Batch` runs many
materials and sources that share the geometry, the observation bins and the
time axis in one persistent kernel; :meth:`Batch.adjoint` replays the same
trajectories and returns dL/d(c, lc, Q) per item, where the coefficients of
variation of item i are ``c * cv / cv[0]`` (``c = cv[0]``), ``lc`` is
``CorrelationLength`` and ``Q`` the intrinsic attenuation (acoustic media).
:class:`EnergyDensity` wraps both as a ``torch.autograd.Function``.
"""
from __future__ import annotations

import copy
import ctypes as C
import math
from typing import List, Optional, Sequence

import numpy as np
import torch

from . import _abi
from .material import MaterialClass, _integral
from .observation import Observation, normalize
from .solver import (DeviceBins, DeviceMaterial, _device, _f64, _get, _stream_ptr, build_geometry, build_source)

__all__ = ["prepare_material", "lc_derivatives", "Batch", "EnergyDensity", "count_scale", "count_scales"]


def prepare_material(base: MaterialClass, d: int, c: Optional[float] = None, lc: Optional[float] = None,
                     Q: Optional[float] = None, knot_threshold: float = 1e-8) -> MaterialClass:
    """A prepared copy of ``base`` with cv scaled to cv[0] = c, lc and Q replaced
    (MaterialClass.prepareSigma, MaterialClass.m:2379-2442).  The default
    tables are the reference's; for gradients use ``knot_threshold=0``
    (tables that follow sigma over the whole angle range)."""
    m = copy.deepcopy(base)
    if c is not None:
        cv = np.asarray(m.coefficients_of_variation, dtype=np.float64).ravel()
        m.coefficients_of_variation = cv * (float(c) / cv[0])
    if lc is not None:
        m.CorrelationLength = float(lc)
    if Q is not None:
        m.Q = float(Q)
    m.Phi = None
    m.sigma = []
    m.invcdf = None
    return MaterialClass.prepareSigma(m, d, knot_threshold=knot_threshold)


def _sigma_fn(base: MaterialClass, lc: float):
    m = copy.deepcopy(base)
    m.CorrelationLength = float(lc)
    m.Phi = None
    m.sigma = []
    m.CalcSigma()
    return m.sigma[0]


def _Sigma(sigma, d: int) -> float:
    # total cross-section as in MaterialClass.prepareSigmaOne (MaterialClass.m:2443-2455)
    if d == 2:
        return 2 * _integral(sigma, 0, np.pi)
    return 2 * np.pi * _integral(lambda th: sigma(th) * np.sin(th), 0, np.pi)


def lc_derivatives(material: MaterialClass, d: int, n: int = 2049, rel: float = 1e-4):
    """d log sigma(theta)/d lc on theta = linspace(0, pi, n), and d log Sigma/d lc,
    by central differences of the acoustic cross-section (MaterialClass.m:204-235)
    at lc (1 +- rel).  Returns (g, dlogSigma)."""
    if not material.acoustics:
        raise ValueError("lc derivatives are implemented for acoustic media")
    lc = float(material.CorrelationLength)
    h = rel * lc
    sp = _sigma_fn(material, lc + h)
    sm = _sigma_fn(material, lc - h)
    th = np.linspace(0.0, np.pi, n)
    a = np.asarray(sp(th), dtype=np.float64)
    b = np.asarray(sm(th), dtype=np.float64)
    ok = (a > 0) & (b > 0)
    g = np.zeros(n)
    g[ok] = (np.log(a[ok]) - np.log(b[ok])) / (2 * h)
    Sp, Sm = _Sigma(sp, d), _Sigma(sm, d)
    dlogS = (math.log(Sp) - math.log(Sm)) / (2 * h)
    return g, dlogS


def count_scale(O: Observation, N: int) -> np.ndarray:
    """Per-bin factor f with energyDensity = count * f, flat in the histogram's
    column-major order (radiativeTransfer.m:160-188)."""
    ones = np.ones(O.size, dtype=np.float64)
    dens, _ = normalize(O, ones, N)
    return np.asarray(dens, dtype=np.float64).reshape(-1, order='F')


def count_scales(O: Observation, Ns: Sequence[int], device) -> torch.Tensor:
    """count_scale of every item (n x O.size) with one host-to-device copy."""
    base = torch.as_tensor(count_scale(O, 1), device=device)
    inv = torch.as_tensor([1.0 / float(N) for N in Ns], dtype=torch.float64, device=device)
    return inv[:, None] * base[None, :]


class Batch:
    """Items (material, source, particle count, seed) sharing geometry and
    observation; owns their device histograms E (n_items x O.size, int32) and
    counters (n_items x 4, int64)."""

    def __init__(self, geometry, O: Observation, materials: Sequence[MaterialClass], sources, Ns: Sequence[int],
                 seeds: Sequence[int], *, device=None, time_steps: int = 0, precision: str = "fp64",
                 offsets: Optional[Sequence[int]] = None, lds_tol: float = 0.0, lds_max_knots: int = 1024):
        n = len(materials)
        if not (len(Ns) == n and len(seeds) == n):
            raise ValueError("materials, Ns and seeds must have the same length")
        if not isinstance(sources, (list, tuple)):
            sources = [sources] * n
        self.dev = _device(device)
        self.d = int(_get(geometry, 'dimension'))
        self.O = O
        self.n = n
        self.materials = list(materials)
        self.Ns = [int(v) for v in Ns]
        self.keep = []
        self.geometry = build_geometry(geometry)
        self.bins = DeviceBins(O, self.dev)
        # prepared MaterialClass objects are uploaded here; anything else carries
        # its rtmc_material in `.c` (native_material.NativeMaterials.items)
        # (lds_tol > 0: with compressed angle tables, which the kernel holds in LDS)
        self.dmats = [DeviceMaterial(m, self.d, self.dev, with_zoeppritz=self.geometry.n_bnd > 0, lds_tol=lds_tol,
                                     lds_max_knots=lds_max_knots)
                      if isinstance(m, MaterialClass) else m for m in materials]
        self._tables = None
        self._native = None
        self.E = torch.zeros((n, O.size), dtype=torch.int32, device=self.dev)
        self.counters = torch.zeros((n, _abi.NCOUNTERS), dtype=torch.int64, device=self.dev)
        self.items = (_abi.BatchItem * max(n, 1))()
        for i in range(n):
            it = self.items[i]
            it.material = self.dmats[i].c
            it.source = build_source(sources[i], self.d, self.keep, self.dev)
            it.n_particles = self.Ns[i]
            it.particle_offset = int(offsets[i]) if offsets is not None else 0
            it.seed = int(seeds[i]) & 0xFFFFFFFFFFFFFFFF
            it.E = self.E[i].data_ptr()
            it.counters = self.counters[i].data_ptr()
        b = _abi.Batch()
        b.geometry = self.geometry
        b.bins = self.bins.c
        b.n_times = int(O.Nt)
        b.time_steps = int(time_steps)
        b.dt = float(np.mean(np.diff(O.t))) if O.Nt > 1 else 0.0
        if precision not in ("fp64", "fp32"):
            raise ValueError("precision must be 'fp64' or 'fp32' (FP32 free flight, FP64 time)")
        b.precision = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_FP32_FLIGHT
        b.n_items = n
        b.items = C.cast(self.items, C.POINTER(_abi.BatchItem))
        self.c = b

    def configure(self, block_size: int = 0, particles_per_thread: int = 0, hist_mode: int = 0,
                  event_batch: int = 0) -> None:
        self.c.block_size = int(block_size)
        self.c.particles_per_thread = int(particles_per_thread)
        self.c.hist_mode = int(hist_mode)
        self.c.event_batch = int(event_batch)

    def forward(self, zero: bool = True):
        """Runs every item; returns (E, counters) (device tensors)."""
        if zero:
            self.E.zero_()
            self.counters.zero_()
        lib = _abi.load()
        _abi.check(lib.rtmc_forward_batched(C.byref(self.c), C.c_void_p(_stream_ptr(self.dev))),
                   "rtmc_forward_batched")
        return self.E, self.counters

    def autotune(self, sample_fraction: float = 0.02, apply: bool = True) -> dict:
        """rtmc_autotune_batched on a sample of every item; applies the winner."""
        t = _abi.Tuning()
        lib = _abi.load()
        _abi.check(lib.rtmc_autotune_batched(C.byref(self.c), float(sample_fraction), C.byref(t),
                                             C.c_void_p(_stream_ptr(self.dev))), "rtmc_autotune_batched")
        out = {"block_size": t.block_size, "particles_per_thread": t.particles_per_thread,
               "hist_mode": t.hist_mode, "event_batch": t.event_batch, "ms": t.ms, "ms_default": t.ms_default}
        if apply:
            self.configure(t.block_size, t.particles_per_thread, t.hist_mode, t.event_batch)
        return out

    @classmethod
    def from_params(cls, geometry, O: Observation, base: MaterialClass, params, sources, Ns: Sequence[int],
                    seeds: Sequence[int], *, knot_threshold: float = 0.0, n_glc: int = 2049, lds_tol: float = 0.0,
                    lds_max_knots: int = 1024, **kw) -> "Batch":
        """A batch of ``base`` at the (c, lc, Q) rows of ``params``, its materials
        and the adjoint's derivative tables prepared natively for all items at
        once (rtmc_prepare_materials, rtmc_lc_derivatives); lds_tol > 0 adds the
        compressed angle tables the kernel holds in LDS."""
        from . import native_material as nm
        d = int(_get(geometry, 'dimension'))
        plist = [nm.params_for(base, d, c, lc, Q, knot_threshold=knot_threshold, lds_tol=lds_tol,
                               lds_max_knots=lds_max_knots) for c, lc, Q in params]
        nat = nm.NativeMaterials(plist)
        B = cls(geometry, O, nat.items, sources, Ns, seeds, **kw)
        B._native = nat  # owns the device tables
        if base.acoustics:
            g, dS = nm.lc_derivatives(plist, n_glc)
            B._tables = [(float(p.cv[0]), float(dS[i]), g[i]) for i, p in enumerate(plist)]
        return B

    def derivative_tables(self, n: int = 2049):
        """(c, dlogSigma_dlc, dlogsigma_dlc table) of every item."""
        if self._tables is not None and all(np.size(t[2]) == n for t in self._tables):
            return self._tables
        out = []
        for m in self.materials:
            g, dS = lc_derivatives(m, self.d, n)
            c = float(np.asarray(m.coefficients_of_variation, dtype=np.float64).ravel()[0])
            out.append((c, dS, g))
        return out

    def adjoint(self, weights: torch.Tensor, tables=None, dhist: bool = False):
        """dL/d(c, lc, Q) of every item (n x 3 float64 device tensor) for the
        per-bin weights dL/dcount (n x O.size); with ``dhist`` also the per-bin
        derivatives d count/d(c, lc, Q) (n x 3 x O.size)."""
        if tables is None:
            tables = self.derivative_tables()
        w = weights.to(device=self.dev, dtype=torch.float32).contiguous()
        if w.shape != (self.n, self.O.size):
            raise ValueError(f"weights must have shape {(self.n, self.O.size)}")
        grad = torch.zeros((self.n, 3), dtype=torch.float64, device=self.dev)
        dh = torch.zeros((self.n, 3, self.O.size), dtype=torch.float64, device=self.dev) if dhist else None
        adj = (_abi.AdjointItem * max(self.n, 1))()
        # one upload for all derivative tables
        ng = max(int(np.size(t[2])) for t in tables)
        host = np.zeros((self.n, ng))
        for i, (_c, _dS, g) in enumerate(tables):
            host[i, :np.size(g)] = g
        gt = _f64(host, self.dev)
        for i, (c, dS, g) in enumerate(tables):
            a = adj[i]
            a.weight = w[i].data_ptr()
            a.grad = grad[i].data_ptr()
            a.dhist = dh[i].data_ptr() if dh is not None else None
            a.cv = float(c)
            a.dlogSigma_dlc = float(dS)
            a.dlogsigma_dlc = gt[i].data_ptr()
            a.n_dlogsigma = int(np.size(g))
        lib = _abi.load()
        _abi.check(lib.rtmc_adjoint_batched(C.byref(self.c), adj, C.c_void_p(_stream_ptr(self.dev))),
                   "rtmc_adjoint_batched")
        torch.cuda.current_stream(self.dev).synchronize()  # keeps `gt` alive until the kernel is done
        return (grad, dh) if dhist else grad

    def energy_density(self) -> torch.Tensor:
        """energyDensity of every item, flat (n x O.size, float64)."""
        return self.E.to(torch.float64) * count_scales(self.O, self.Ns, self.dev)


class EnergyDensity(torch.autograd.Function):
    """energyDensity (n x O.size) of n acoustic materials as a function of
    their (c, lc, Q) (an n x 3 float64 tensor), with the adjoint kernel as
    backward.  ``spec`` = dict(base=MaterialClass, geometry, O, sources, Ns,
    seeds, device, time_steps, precision, and optionally knot_threshold,
    tuning, lds_tol)."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, spec: dict):
        d = int(_get(spec["geometry"], 'dimension'))
        P = params.detach().cpu().numpy().astype(np.float64)
        kt = spec.get("knot_threshold", 0.0)
        mats = [prepare_material(spec["base"], d, c=p[0], lc=p[1], Q=p[2], knot_threshold=kt) for p in P]
        B = Batch(spec["geometry"], spec["O"], mats, spec["sources"], spec["Ns"], spec["seeds"],
                  device=spec.get("device"), time_steps=spec.get("time_steps", 0),
                  precision=spec.get("precision", "fp64"), lds_tol=spec.get("lds_tol", 0.0))
        if spec.get("tuning"):
            B.configure(**spec["tuning"])
        B.forward()
        dens = B.energy_density()
        ctx.B = B
        return dens

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        B = ctx.B
        w = grad_out.to(B.dev, torch.float64) * count_scales(B.O, B.Ns, B.dev)
        g = B.adjoint(w)
        return g.to(torch.float64).cpu(), None

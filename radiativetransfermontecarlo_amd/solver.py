"""radiativeTransfer and the per-function API, driving librtmc on the GPU.

``radiativeTransfer(geometry, source, material, observation)`` keeps the
reference's signature and ``obs`` fields (radiativeTransfer.m:1-207).  The
packet x time loop runs on the GPU either as one fused kernel
(``engine='fused'``, default) or as the reference's sequence of function
calls (``engine='functions'``: initializeParticle, then per observation time
propagateParticleSmallDt / propagateParticle and observeTime, each one a
kernel on a packet of particles in device memory).

Device memory and streams come from PyTorch (plumbing only); all particle
arithmetic is in librtmc.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import time as _time
import warnings
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _abi
from .material import MaterialClass, InverseCDF
from .observation import Observation, initializeObservation, normalize

__all__ = ["radiativeTransfer", "Particles", "Draws", "DeviceMaterial", "initializeParticle",
           "propagateParticleSmallDt", "propagateParticle", "observeTime", "build_source",
           "build_geometry", "DeviceBins", "forward_histogram"]

NPK = 100_000   # particles per packet (radiativeTransfer.m:23)


def _get(obj, name, default=None):
    if isinstance(obj, dict):
        return obj.get(name, default)
    return getattr(obj, name, default)


def _device(device=None) -> torch.device:
    if device is not None:
        return torch.device(device)
    if not torch.cuda.is_available():
        raise _abi.RTMCError("librtmc needs a GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def _stream_ptr(device: torch.device) -> int:
    return int(torch.cuda.current_stream(device).cuda_stream)


def _f64(a, device) -> torch.Tensor:
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)


# ----------------------------------------------------------------------
# geometry / source descriptors

def build_geometry(geometry) -> _abi.Geometry:
    g = _abi.Geometry()
    g.dimension = int(_get(geometry, 'dimension'))
    frame = _get(geometry, 'frame', 'spherical')
    if frame not in _abi.FRAMES:
        raise ValueError('unknown frame type')
    g.frame = _abi.FRAMES[frame]
    bnd = _get(geometry, 'bnd', None)
    if bnd is None:
        bnd = []
    if isinstance(bnd, dict):
        bnd = [bnd]
    if len(bnd) > _abi.MAX_BND:
        raise ValueError(f"at most {_abi.MAX_BND} boundaries are supported")
    g.n_bnd = len(bnd)
    for i, b in enumerate(bnd):
        typ = _get(b, 'type', None)
        typ = 'reflective' if (typ is None or typ == '') else str(typ).lower()   # propagateParticleSmallDt.m:153-157
        if typ not in ('reflective', 'absorbing'):
            raise ValueError(f'Unknown boundary type "{typ}". Use "reflective" or "absorbing".')
        g.bnd[i].dir = int(_get(b, 'dir'))
        g.bnd[i].type = _abi.BND_REFLECTIVE if typ == 'reflective' else _abi.BND_ABSORBING
        g.bnd[i].val = float(_get(b, 'val'))
    return g


def _null_basis(n: np.ndarray):
    """Orthonormal basis of null(n) for a row vector n, from the SVD as
    MATLAB's null() computes it (initializeParticle.m:98-100)."""
    _, _, vh = np.linalg.svd(n.reshape(1, 3))
    return vh[1], vh[2]


def build_source(source, d: int, keep: list, device) -> _abi.Source:
    """rtmc_source from the MATLAB source struct (initializeParticle.m:3-147)."""
    s = _abi.Source()
    typ = _get(source, 'type', 'point') or 'point'
    pos = np.asarray(_get(source, 'position', [0, 0, 0]), dtype=np.float64).ravel()
    if d == 2 and pos.size == 2:
        pos = np.concatenate([pos, [0.0]])
    s.position[:] = pos[:3]
    s.lambda_ = float(_get(source, 'lambda', 0.0) or 0.0)
    pol = _get(source, 'polarization', 'P')
    s.polarization = 1 if pol == 'S' else 0
    if typ == 'point':
        s.type = _abi.SRC_POINT
        dirn = _get(source, 'direction', 'uniform')
        s.direction = {'outgoing': _abi.DIR_OUTGOING, 'upper': _abi.DIR_UPPER}.get(dirn, _abi.DIR_UNIFORM) \
            if isinstance(dirn, str) else _abi.DIR_UNIFORM
        radial = _get(source, 'radial', None)
        if radial is not None:
            knots = _radial_inverse_cdf(radial, d)
            tx = _f64(knots.x, device)
            ty = _f64(knots.y, device)
            keep += [tx, ty]
            s.radial.n = knots.x.size
            s.radial.x = tx.data_ptr()
            s.radial.y = ty.data_ptr()
            s.radial.extrap = _abi.EXTRAP_NEAREST
    elif typ == 'plane':
        if d != 3:
            raise ValueError("plane sources are only available in 3D")
        s.type = _abi.SRC_PLANE
        ap = (_get(source, 'aperture', 'rectangle') or 'rectangle').lower()
        if ap not in ('rectangle', 'circle'):
            raise ValueError("Unknown plane-wave aperture: use 'rectangle' or 'circle'.")
        s.aperture = _abi.APERTURE_RECTANGLE if ap == 'rectangle' else _abi.APERTURE_CIRCLE
        dirn = _get(source, 'direction')
        dv = np.asarray(dirn, dtype=np.float64).ravel()
        if dv.size == 1:
            axes = {1: [1, 0, 0], 2: [0, 1, 0], 3: [0, 0, 1], -1: [-1, 0, 0], -2: [0, -1, 0], -3: [0, 0, -1]}
            if int(dv[0]) not in axes or dv[0] != int(dv[0]):
                raise ValueError('For a scalar plane-source direction, use ±1, ±2, or ±3.')
            n = np.asarray(axes[int(dv[0])], dtype=np.float64)
        elif dv.size == 3:
            n = dv
        else:
            raise ValueError('For a plane source, source.direction must be a 3-component vector or a '
                             'scalar axis ±1, ±2, ±3.')
        n = n / np.linalg.norm(n)
        u, v = _null_basis(n)
        s.normal[:] = n
        s.basis_u[:] = u
        s.basis_v[:] = v
        ext = np.asarray(_get(source, 'extent'), dtype=np.float64).ravel()
        if s.aperture == _abi.APERTURE_RECTANGLE:
            s.extent[0] = ext[0]
            s.extent[1] = ext[0] if ext.size == 1 else ext[1]
        else:
            if ext.size != 1:
                raise ValueError("For source.aperture = 'circle', source.extent must be a scalar radius.")
            s.extent[0] = ext[0]
    else:
        raise ValueError(f"unknown source type {typ!r}")
    return s


def _radial_inverse_cdf(f, d: int) -> InverseCDF:
    """Local inverseCDF of initializeParticle.m:175-203 for a user radial law.
    ``f`` is an InverseCDF-like table (``GridVectors``/``Values``) or a callable
    with an ``Rmax`` attribute."""
    if hasattr(f, 'GridVectors'):
        Rmax = float(f.GridVectors[0][-1])
        fun = f
    else:
        Rmax = float(getattr(f, 'Rmax'))
        fun = f
    Nth = 10000
    xth = np.linspace(0, Rmax, Nth)
    g = xth ** (d - 1) * np.asarray(fun(xth), dtype=np.float64)
    intF = np.trapezoid(g, xth) if hasattr(np, 'trapezoid') else np.trapz(g, xth)
    pdf = g / intF
    cdf = np.cumsum(pdf) * np.mean(np.diff(xth))
    cdf[0] = 0.0
    cdf[-1] = 1.0
    ind = np.nonzero(np.diff(cdf) > 0)[0]
    ind = np.unique(np.concatenate([ind, ind + 1]))
    uc, idx = np.unique(np.sort(cdf[ind]), return_index=True)
    ds = np.sort(xth[ind])
    ux = ds[idx]
    return InverseCDF(uc, ux, extrap="nearest")


# ----------------------------------------------------------------------
# material on the device

class DeviceMaterial:
    """A prepared MaterialClass with its tables uploaded (rtmc_material).
    lds_tol > 0 also uploads compressed copies of the angle tables
    (rtmc_compress_table with this tolerance in radians and at most
    lds_max_knots knots each), which the tracing kernel then holds in LDS."""

    def __init__(self, material: MaterialClass, d: int, device, with_zoeppritz: bool = False,
                 lds_tol: float = 0.0, lds_max_knots: int = 1024):
        self.keep = []
        m = _abi.Material()
        m.acoustics = 1 if material.acoustics else 0
        if material.acoustics:
            tabs = [material.invcdf]
            m.vp = float(material.vp if material.vp is not None else material.v)
            m.vs = 0.0
            mft = float(np.asarray(material.meanFreeTime).ravel()[0])
            m.mft[0] = mft
            m.mft[1] = mft
            m.P2P = 1.0
            m.S2S = 1.0
        else:
            inv = material.invcdf
            tabs = [inv[0][0], inv[0][1], inv[1][0], inv[1][1]]
            m.vp = float(material.vp)
            m.vs = float(material.vs)
            mft = np.asarray(material.meanFreeTime, dtype=np.float64).ravel()
            m.mft[0] = mft[0]
            m.mft[1] = mft[1] if mft.size > 1 else mft[0]
            m.P2P = float(material.P2P)
            m.S2S = float(material.S2S)
        Q = np.asarray(material.Q if material.Q is not None else math.inf, dtype=np.float64).ravel()
        if Q.size == 0:
            Q = np.array([math.inf])
        m.QP = Q[0]
        m.QS = Q[1] if (Q.size > 1 and not material.acoustics) else Q[0]
        fr = material.Frequency
        m.frequency = float(np.asarray(fr).ravel()[0]) if fr is not None and np.size(fr) > 0 else math.nan
        for k, t in enumerate(tabs):
            x, y, ex = _table_arrays(t)
            tx = _f64(x, device)
            ty = _f64(y, device)
            self.keep += [tx, ty]
            m.invcdf[k].n = x.size
            m.invcdf[k].x = tx.data_ptr()
            m.invcdf[k].y = ty.data_ptr()
            m.invcdf[k].extrap = ex
            if lds_tol > 0.0:
                from .native_material import compress_table
                kn, used = compress_table(x, y, lds_tol, lds_max_knots)
                tk = torch.from_numpy(kn.view(np.uint8).copy()).to(device)
                self.keep.append(tk)
                m.lds[k].m = kn.size
                m.lds[k].extrap = ex
                m.lds[k].tol = used
                m.lds[k].knots = tk.data_ptr()
        if with_zoeppritz and not material.acoustics:
            out = MaterialClass.Zoeppritz(material)
            deg = _f64(out['j1_deg'], device)
            rpp = _f64(out['E_Rpp'], device)
            rss = _f64(out['E_Rsvsv'], device)
            self.keep += [deg, rpp, rss]
            m.n_zoe = deg.numel()
            m.zoe_deg = deg.data_ptr()
            m.zoe_Rpp = rpp.data_ptr()
            m.zoe_Rsvsv = rss.data_ptr()
        self.c = m


def _table_arrays(t):
    if isinstance(t, InverseCDF):
        return t.x, t.y, (_abi.EXTRAP_NEAREST if t.extrap == 'nearest' else _abi.EXTRAP_LINEAR)
    if hasattr(t, 'GridVectors'):
        return (np.asarray(t.GridVectors[0], dtype=np.float64), np.asarray(t.Values, dtype=np.float64),
                _abi.EXTRAP_LINEAR)
    if isinstance(t, tuple) and len(t) == 2:
        return np.asarray(t[0], dtype=np.float64), np.asarray(t[1], dtype=np.float64), _abi.EXTRAP_LINEAR
    raise TypeError("invcdf entries must be griddedInterpolant-like tables")


class DeviceBins:
    """rtmc_bins for an Observation, with the edges uploaded."""

    def __init__(self, O: Observation, device):
        b = _abi.Bins()
        self.keep = []
        for k in range(4):
            e = _f64(O.edges[k], device)
            self.keep.append(e)
            b.n_edges[k] = int(O.edges[k].size)
            b.edges[k] = e.data_ptr()
            b.stride[k] = int(O.strides[k])
        b.time_stride = int(O.time_stride)
        b.mode_stride = int(O.mode_stride)
        b.direction_variable = _abi.DIRVARS[getattr(O, 'direction_variable', 'dir_z')]
        self.c = b


# ----------------------------------------------------------------------
# per-function API on a packet of particles

@dataclass
class Particles:
    """The particle struct P (initializeParticle.m:163-170) in device memory.
    x, dir, perp have shape (3, N): row k is column k of the MATLAB N x 3 array."""
    d: int
    N: int
    x: torch.Tensor
    dir: torch.Tensor
    perp: torch.Tensor
    p: torch.Tensor
    alive: torch.Tensor
    t: torch.Tensor

    @staticmethod
    def empty(N: int, d: int, device) -> "Particles":
        f = dict(dtype=torch.float64, device=device)
        return Particles(d, N, torch.zeros((3, N), **f), torch.zeros((3, N), **f), torch.zeros((3, N), **f),
                         torch.zeros(N, dtype=torch.uint8, device=device),
                         torch.zeros(N, dtype=torch.uint8, device=device), torch.zeros(N, **f))

    def c(self) -> _abi.Particles:
        P = _abi.Particles()
        P.n = self.N
        P.dimension = self.d
        P.x = self.x.data_ptr()
        P.dir = self.dir.data_ptr()
        P.perp = self.perp.data_ptr()
        P.p = self.p.data_ptr()
        P.alive = self.alive.data_ptr()
        P.t = self.t.data_ptr()
        return P

    def to_numpy(self) -> dict:
        return dict(x=self.x.cpu().numpy().T.copy(), dir=self.dir.cpu().numpy().T.copy(),
                    perp=self.perp.cpu().numpy().T.copy(), p=self.p.cpu().numpy().astype(bool),
                    alive=self.alive.cpu().numpy().astype(bool), t=self.t.cpu().numpy().copy())


class Draws:
    """Random variates of one call: recorded (replay) or Philox."""

    def __init__(self, kind: int, replay: Optional[torch.Tensor] = None, seed: int = 0, stream: int = 0):
        self.kind = kind
        self.replay = replay
        self.seed = int(seed)
        self.stream = int(stream)

    @staticmethod
    def philox(seed: int, stream: int = 0) -> "Draws":
        return Draws(_abi.DRAWS_PHILOX, seed=seed, stream=stream)

    @staticmethod
    def recorded(values: torch.Tensor) -> "Draws":
        """values: float64 device tensor of shape (steps, slots, N)."""
        if values.dtype != torch.float64 or values.dim() != 3 or not values.is_contiguous():
            raise ValueError("replay draws must be a contiguous float64 (steps, slots, N) tensor")
        return Draws(_abi.DRAWS_REPLAY, replay=values)

    def c(self) -> _abi.Draws:
        D = _abi.Draws()
        D.kind = self.kind
        D.seed = self.seed & 0xFFFFFFFFFFFFFFFF
        D.stream = self.stream & 0xFFFFFFFFFFFFFFFF
        if self.kind == _abi.DRAWS_REPLAY:
            D.replay = self.replay.data_ptr()
            D.n_steps = self.replay.shape[0]
            D.n_slots = self.replay.shape[1]
        return D


def initializeParticle(N: int, d: int, acoustics: bool, source, draws: Draws, device=None) -> Particles:
    """initializeParticle.m:1-171 on the GPU."""
    dev = _device(device)
    keep = []
    s = build_source(source, d, keep, dev)
    P = Particles.empty(int(N), int(d), dev)
    Pc = P.c()
    Dc = draws.c()
    lib = _abi.load()
    _abi.check(lib.rtmc_init_particles(C.byref(s), 1 if acoustics else 0, C.byref(Dc), C.byref(Pc),
                                       C.c_void_p(_stream_ptr(dev))), "rtmc_init_particles")
    return P


def propagateParticleSmallDt(mat: DeviceMaterial, geometry, P: Particles, T: float, draws: Draws,
                             T_eff: Optional[float] = None) -> _abi.SubstepInfo:
    """propagateParticleSmallDt.m:1-647 on the GPU (updates P in place)."""
    g = build_geometry(geometry) if not isinstance(geometry, _abi.Geometry) else geometry
    info = _abi.SubstepInfo()
    Pc = P.c()
    Dc = draws.c()
    lib = _abi.load()
    _abi.check(lib.rtmc_propagate_smalldt(C.byref(mat.c), C.byref(g), C.byref(Pc), float(T),
                                          math.nan if T_eff is None else float(T_eff), C.byref(Dc),
                                          C.byref(info), C.c_void_p(_stream_ptr(P.x.device))),
               "rtmc_propagate_smalldt")
    return info


def propagateParticle(mat: DeviceMaterial, P: Particles, T: float, draws: Draws) -> None:
    """propagateParticle.m:1-96 on the GPU (updates P in place)."""
    Pc = P.c()
    Dc = draws.c()
    lib = _abi.load()
    _abi.check(lib.rtmc_propagate_events(C.byref(mat.c), C.byref(Pc), float(T), C.byref(Dc),
                                         C.c_void_p(_stream_ptr(P.x.device))), "rtmc_propagate_events")


def observeTime(geometry, bins: DeviceBins, acoustics: bool, P: Particles, E: torch.Tensor, it: int) -> None:
    """E(:,:,it,:) += observeTime(...) (observeTime.m:1-146, radiativeTransfer.m:75-76).
    ``it`` is the 0-based time index; E is the flat int32 device histogram."""
    g = build_geometry(geometry) if not isinstance(geometry, _abi.Geometry) else geometry
    Pc = P.c()
    lib = _abi.load()
    _abi.check(lib.rtmc_observe(C.byref(g), C.byref(bins.c), 1 if acoustics else 0, C.byref(Pc),
                                C.c_void_p(E.data_ptr()), int(it), C.c_void_p(_stream_ptr(P.x.device))),
               "rtmc_observe")


# ----------------------------------------------------------------------
# whole solve

def _default_seed() -> int:
    return int.from_bytes(np.random.default_rng().bytes(8), "little") & 0x7FFFFFFFFFFFFFFF


def forward_histogram(geometry, source, material: MaterialClass, O: Observation, N: int, *, seed: int,
                      device, time_steps: int = 0, particle_offset: int = 0, E: Optional[torch.Tensor] = None,
                      counters: Optional[torch.Tensor] = None, block_size: int = 0, hist_mode: int = 0,
                      particles_per_thread: int = 0, dmat: Optional[DeviceMaterial] = None,
                      precision: str = "fp64", bins: Optional[DeviceBins] = None):
    """One rtmc_forward call: adds the counts of N particles into E (flat int32
    device tensor, allocated when None).  Returns (E, counters, keep-alive)."""
    d = int(_get(geometry, 'dimension'))
    dev = _device(device)
    keep = []
    prob = _abi.Problem()
    prob.geometry = build_geometry(geometry)
    prob.source = build_source(source, d, keep, dev)
    if dmat is None:
        dmat = DeviceMaterial(material, d, dev, with_zoeppritz=prob.geometry.n_bnd > 0)
    prob.material = dmat.c
    if bins is None:
        bins = DeviceBins(O, dev)
    prob.bins = bins.c
    prob.n_particles = int(N)
    prob.particle_offset = int(particle_offset)
    prob.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    prob.n_times = int(O.Nt)
    prob.time_steps = int(time_steps)
    prob.dt = float(np.mean(np.diff(O.t))) if O.Nt > 1 else 0.0
    if precision not in ("fp64", "fp32"):
        raise ValueError("precision must be 'fp64' or 'fp32' (FP32 free flight, FP64 time)")
    prob.precision = _abi.PREC_FP64 if precision == "fp64" else _abi.PREC_FP32_FLIGHT
    prob.block_size = int(block_size)
    prob.particles_per_thread = int(particles_per_thread)
    prob.hist_mode = int(hist_mode)
    if E is None:
        E = torch.zeros(O.size, dtype=torch.int32, device=dev)
    if counters is None:
        counters = torch.zeros(_abi.NCOUNTERS, dtype=torch.int64, device=dev)
    lib = _abi.load()
    _abi.check(lib.rtmc_forward(C.byref(prob), C.c_void_p(E.data_ptr()), C.c_void_p(counters.data_ptr()),
                                C.c_void_p(_stream_ptr(dev))), "rtmc_forward")
    return E, counters, (keep, bins, dmat)


def radiativeTransfer(geometry, source, material: MaterialClass, observation, *, seed: Optional[int] = None,
                      engine: str = "fused", device=None, allow_nd: bool = False, return_histogram: bool = False,
                      precision: str = "fp64", lds_tol: float = 0.0):
    """obs = radiativeTransfer(geometry, source, material, observation)
    (radiativeTransfer.m:1-207) on one GPU.

    engine='fused': one persistent kernel for all packets (Philox draws).
    engine='functions': the reference's packet x time loop of function calls.
    precision='fp32' (fused engine): FP32 direction/flight arithmetic, FP64 time.
    lds_tol > 0 (fused engine): scattering angles from compressed copies of the
    angle tables held in LDS, within lds_tol radians of the full tables.
    """
    d = int(_get(geometry, 'dimension'))
    if isinstance(geometry, dict) and 'frame' not in geometry:
        geometry = dict(geometry, frame='spherical')                     # :7-9
    acoustics = bool(material.acoustics)
    Q = material.Q
    if Q is not None and np.size(Q) == 1 and not acoustics:              # :12-17
        warnings.warn(f"Only one quality factor was provided for an elastic simulation. The same value "
                      f"Q = {float(np.asarray(Q).ravel()[0]):g} will be used for both P and S waves.")
    Nreq = int(_get(source, 'numberParticles'))
    Np = int(math.ceil(Nreq / NPK))                                      # :23-24
    N = Np * NPK
    O = initializeObservation(geometry, acoustics, observation, N, allow_nd=allow_nd)   # :27-28
    material = MaterialClass.prepareSigma(material, d)                   # :31
    dev = _device(device)
    if seed is None:
        seed = _default_seed()
    time_steps = int(material.timeSteps)
    bnd = _get(geometry, 'bnd', None)
    if time_steps == 1 and bnd:
        warnings.warn("material.timeSteps = 1 (propagateParticle) ignores geometry.bnd; boundaries dropped")
        geometry = dict(geometry)
        geometry.pop('bnd')
    if engine == "fused":
        dmat = DeviceMaterial(material, d, dev, with_zoeppritz=bool(_get(geometry, 'bnd', None)), lds_tol=lds_tol)
        E, counters, _k = forward_histogram(geometry, source, material, O, N, seed=seed, device=dev,
                                            time_steps=time_steps, dmat=dmat, precision=precision)
    elif engine == "functions":
        if lds_tol > 0.0:
            raise ValueError("lds_tol applies to the fused engine (the function-level kernels read the full tables)")
        E = _functions_engine(geometry, source, material, O, N, Np, seed, dev, time_steps)
    else:
        raise ValueError(f"unknown engine {engine!r}")
    Eh = E.cpu().numpy().view(np.uint32)
    dens, energy = normalize(O, Eh, N)                                   # :160-191
    obs = dict(O.obs)
    obs['energyDensity'] = dens
    obs['energy'] = energy
    if return_histogram:
        obs['E'] = Eh.reshape(O.shape, order='F')
    return obs


def _functions_engine(geometry, source, material, O: Observation, N: int, Np: int, seed: int, dev,
                      time_steps: int) -> torch.Tensor:
    d = int(_get(geometry, 'dimension'))
    acoustics = bool(material.acoustics)
    g = build_geometry(geometry)
    dmat = DeviceMaterial(material, d, dev, with_zoeppritz=g.n_bnd > 0)
    bins = DeviceBins(O, dev)
    E = torch.zeros(O.size, dtype=torch.int32, device=dev)
    dt = float(np.mean(np.diff(O.t))) if O.Nt > 1 else 0.0            # :37
    for ip in range(Np):                                                 # :105
        key = (seed + 0x9E3779B97F4A7C15 * (ip + 1)) & 0xFFFFFFFFFFFFFFFF
        P = initializeParticle(NPK, d, acoustics, source, Draws.philox(key, 0), device=dev)
        for it in range(1, O.Nt):                                        # :114
            t_val = 0.0 + it * dt                                        # :117
            dr = Draws.philox(key, it)
            if time_steps == 0:
                propagateParticleSmallDt(dmat, g, P, t_val, dr)
            else:
                propagateParticle(dmat, P, t_val, dr)
            observeTime(g, bins, acoustics, P, E, it)
    return E

"""Observation grid and normalisation of the energy histograms.

Restates ``initializeObservation.m`` (bin edges, centres, widths, volume
weights, the choice of the two histogram variables ``ibins``) and the
normalisation at the end of ``radiativeTransfer.m`` (:160-191).

Beyond the reference, ``allow_nd=True`` accepts more than two binned
variables (the reference raises, initializeObservation.m:93-95): the
histogram then has one axis per binned variable, in the order x, y, z,
directions, and is normalised by the same rule.
"""
from __future__ import annotations

import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import _abi

__all__ = ["Observation", "initializeObservation", "normalize"]


def _get(obj, name, default=None):
    if isinstance(obj, dict):
        return obj.get(name, default)
    return getattr(obj, name, default)


def _volume_spherical(d, r, d0r):
    """volumeEnergySpherical (initializeObservation.m:173-187)."""
    if d == 2:
        dr = r * d0r
        if r.size and r[0] == 0:
            dr[0] = d0r[0] ** 2 / 8
    else:
        dr = r ** 2 * d0r + d0r ** 3 / 12
        if r.size and r[0] == 0:
            dr[0] = d0r[0] ** 3 / 24
    return dr


def _volume_cylindrical(r, d0r):
    """volumeEnergyCylindrical (initializeObservation.m:188-193)."""
    dr = r * d0r
    if r.size and r[0] == 0:
        dr[0] = d0r[0] ** 2 / 8
    return dr


@dataclass
class Observation:
    """The ``obs`` structure (initializeObservation.m:136-156) plus the
    histogram layout used by the kernels."""
    obs: Dict
    ibins: List[int]                 # 1-based binned variables (x=1, y=2, z=3, directions=4)
    edges: List[np.ndarray]          # per variable x, y, z, directions (z unused in 2-D)
    strides: List[int]               # histogram stride of each variable (0 = one bin)
    shape: tuple                     # histogram shape (bins..., Nt, nmode), column-major
    time_stride: int
    mode_stride: int
    n_modes: int
    t: np.ndarray
    Nt: int
    widths: List[np.ndarray] = field(default_factory=list)   # dx (volume weight), dy, dz, dpsi
    direction_variable: str = "dir_z"  # 'dir_z' (observeTime.m:24-35) or 'position' (README.md:209)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape))


def initializeObservation(geometry, acoustics: bool, observation, N: int, allow_nd: bool = False) -> Observation:
    d = int(_get(geometry, 'dimension'))
    frame = _get(geometry, 'frame', 'spherical')
    time = np.atleast_1d(np.asarray(_get(observation, 'time'), dtype=np.float64))
    t = np.concatenate([[0.0], np.setdiff1d(time, [0.0])])               # :9-10
    Nt = t.size
    binX = np.atleast_1d(np.asarray(_get(observation, 'x'), dtype=np.float64))
    x = (binX[:-1] + binX[1:]) / 2
    Nx = x.size
    dx = np.diff(binX)
    if frame == 'cartesian':                                              # :30-41
        binY = np.array([-np.inf, np.inf])
        binZ = np.array([-np.inf, np.inf])
    elif frame == 'cylindrical':
        binY = np.array([-np.pi, np.pi])
        binZ = np.array([-np.inf, np.inf])
    elif frame == 'spherical':
        binY = np.array([-np.pi, np.pi])
        binZ = np.array([-np.pi / 2, np.pi / 2])
    else:
        raise ValueError('unknown frame type')
    oy = _get(observation, 'y')
    if oy is not None and np.size(oy) > 0:
        binY = np.atleast_1d(np.asarray(oy, dtype=np.float64))
    with np.errstate(invalid='ignore'):
        y = (binY[:-1] + binY[1:]) / 2
    Ny = y.size
    dy = np.diff(binY)
    oz = _get(observation, 'z')
    if d == 3 and oz is not None and np.size(oz) > 0:
        binZ = np.atleast_1d(np.asarray(oz, dtype=np.float64))
    if d == 3 and frame == 'spherical':
        binZ = np.sin(binZ)                                              # :56-58
    with np.errstate(invalid='ignore'):
        z = (binZ[:-1] + binZ[1:]) / 2
    Nz = z.size
    dz = np.diff(binZ)
    od = _get(observation, 'directions')
    if od is not None and np.size(od) > 0:
        binPsi = np.atleast_1d(np.asarray(od, dtype=np.float64))
    else:
        binPsi = np.array([0.0, np.pi])
    if d == 3:
        binPsi = np.cos(binPsi[::-1])                                    # :71-73
    psi = (binPsi[:-1] + binPsi[1:]) / 2
    Npsi = psi.size
    dpsi = np.diff(binPsi)
    if frame == 'spherical':                                             # :79-83
        dx = _volume_spherical(d, x, dx)
    elif frame == 'cylindrical':
        dx = _volume_cylindrical(x, dx)

    counts = [Nx, Ny, Nz, Npsi]
    ibins = [k + 1 for k in range(4) if counts[k] > 1]                   # :86
    if d == 2:
        ibins = [k for k in ibins if k != 3]
    if len(ibins) == 0:                                                  # :89-98
        ibins = [1, 2]
    elif len(ibins) == 1:
        ibins = [1, ibins[0]] if ibins[0] != 1 else [1, 2]
    if len(ibins) > 2 and not allow_nd:
        raise ValueError('histograms can only be constructed along two directions')
    if len(ibins) == 2 and ibins not in ([1, 2], [1, 3], [1, 4], [2, 4], [3, 4], [2, 3]):
        raise ValueError(f"Bin combination {ibins} not implemented yet")

    n_modes = 1 if acoustics else 2
    strides = [0, 0, 0, 0]
    shape = []
    s = 1
    for k in ibins:
        strides[k - 1] = s
        shape.append(counts[k - 1])
        s *= counts[k - 1]
    time_stride = s
    mode_stride = s * Nt
    shape = tuple(shape) + (Nt, n_modes)

    obs = dict(d=d, acoustics=bool(acoustics), N=int(N), t=t, Nt=Nt, x=x, Nx=Nx, dx=dx, binX=binX,
               y=y, Ny=Ny, dy=dy, binY=binY, z=z, Nz=Nz, dz=dz, binZ=binZ,
               psi=psi, Npsi=Npsi, dpsi=dpsi, binPsi=binPsi)
    edges = [binX, binY, binZ, binPsi]
    # what `directions` measures: the reference's code bins dir_z (observeTime.m:24-35);
    # README.md:209 describes the angle between the direction and the position vector,
    # available as observation['direction_variable'] = 'position'
    dv = _get(observation, 'direction_variable', 'dir_z') or 'dir_z'
    if dv not in ('dir_z', 'position'):
        raise ValueError("observation.direction_variable must be 'dir_z' or 'position'")
    return Observation(obs=obs, ibins=ibins, edges=edges, strides=strides, shape=shape,
                       time_stride=time_stride, mode_stride=mode_stride, n_modes=n_modes, t=t, Nt=Nt,
                       widths=[dx, dy, dz, dpsi], direction_variable=dv)


def normalize(O: Observation, E: np.ndarray, N: int):
    """energyDensity and energy from the uint32 histogram E (radiativeTransfer.m:160-191).

    energyDensity = E / (N * dx * dy [* dz in 3-D]): every spatial variable is
    divided by its width (volume weight for x in spherical and cylindrical
    frames), the direction variable is not -- this is what each ``ibins``
    branch of :161-188 reduces to.  ``energy`` sums over the bin axes.
    """
    d = O.obs['d']
    Ef = np.asarray(E).reshape(O.shape, order='F').astype(np.float64)
    dens = Ef / float(N)
    nb = len(O.ibins)
    spatial = [0, 1, 2] if d == 3 else [0, 1]
    for v in spatial:
        w = np.asarray(O.widths[v], dtype=np.float64)
        if (v + 1) in O.ibins:
            ax = O.ibins.index(v + 1)
            sh = [1] * Ef.ndim
            sh[ax] = w.size
            dens = dens / w.reshape(sh)
        else:
            dens = dens / float(w[0])
    energy = Ef.sum(axis=tuple(range(nb))) / float(N)                  # (Nt, nmode)
    if O.n_modes == 1:
        dens = dens[..., 0]
        energy = energy[:, 0]
    return dens, energy

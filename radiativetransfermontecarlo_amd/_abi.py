"""ctypes mirror of include/rtmc.h and the loader of librtmc.so.

The library is built in-tree (``__graft_entry__.build()``) next to this file.
Loading fails loudly when it is missing: there is no CPU fallback for the
particle kernels.
"""
from __future__ import annotations

import ctypes as C
import os

ABI_VERSION = 6

OK = 0
ERR_ARG = -1
ERR_HIP = -2
ERR_DRAWS = -3
ERR_NOMEM = -4
ERR_TABLE = -5

MAX_BND = 8

FRAME_CARTESIAN = 0
FRAME_CYLINDRICAL = 1
FRAME_SPHERICAL = 2
FRAMES = {"cartesian": FRAME_CARTESIAN, "cylindrical": FRAME_CYLINDRICAL, "spherical": FRAME_SPHERICAL}

BND_REFLECTIVE = 0
BND_ABSORBING = 1

SRC_POINT = 0
SRC_PLANE = 1
DIR_UNIFORM = 0
DIR_OUTGOING = 1
DIR_UPPER = 2
APERTURE_RECTANGLE = 0
APERTURE_CIRCLE = 1

EXTRAP_LINEAR = 0
EXTRAP_NEAREST = 1

DRAWS_REPLAY = 0
DRAWS_PHILOX = 1

PREC_FP64 = 0
PREC_FP32_FLIGHT = 1

DIRVAR_DIRZ = 0
DIRVAR_POSITION = 1
DIRVARS = {"dir_z": DIRVAR_DIRZ, "position": DIRVAR_POSITION}

HIST_AUTO = 0
HIST_GLOBAL = 1
HIST_LDS = 2
HIST_TMAJOR = 3

# device scratch of rtmc_forward_noalloc per GPU
NOALLOC_SCRATCH_BYTES = 64 << 20

CNT_SCATTER = 0
CNT_DEPOSIT = 1
CNT_BOUNDARY = 2
CNT_KILLED = 3
NCOUNTERS = 4

# draw-slot counts of the entry points (include/rtmc.h)
INIT_SLOTS = 8
EVENT_SLOTS = 6


def smalldt_slots(n_bnd: int) -> int:
    return 6 + 2 * n_bnd


class Boundary(C.Structure):
    _fields_ = [("dir", C.c_int32), ("type", C.c_int32), ("val", C.c_double)]


class Geometry(C.Structure):
    _fields_ = [
        ("dimension", C.c_int32),
        ("frame", C.c_int32),
        ("n_bnd", C.c_int32),
        ("pad_", C.c_int32),
        ("bnd", Boundary * MAX_BND),
    ]


class Table(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("x", C.c_void_p),
        ("y", C.c_void_p),
        ("extrap", C.c_int32),
        ("pad_", C.c_int32),
    ]


class LdsKnot(C.Structure):
    _fields_ = [("u", C.c_double), ("theta", C.c_float), ("slope", C.c_float)]


class LdsTable(C.Structure):
    _fields_ = [("m", C.c_int32), ("extrap", C.c_int32), ("tol", C.c_double), ("knots", C.c_void_p)]


class Material(C.Structure):
    _fields_ = [
        ("acoustics", C.c_int32),
        ("pad_", C.c_int32),
        ("vp", C.c_double),
        ("vs", C.c_double),
        ("mft", C.c_double * 2),
        ("P2P", C.c_double),
        ("S2S", C.c_double),
        ("QP", C.c_double),
        ("QS", C.c_double),
        ("frequency", C.c_double),
        ("invcdf", Table * 4),
        ("n_zoe", C.c_int64),
        ("zoe_deg", C.c_void_p),
        ("zoe_Rpp", C.c_void_p),
        ("zoe_Rsvsv", C.c_void_p),
        ("lds", LdsTable * 4),
    ]


class Source(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("direction", C.c_int32),
        ("aperture", C.c_int32),
        ("polarization", C.c_int32),
        ("position", C.c_double * 3),
        ("lambda_", C.c_double),
        ("normal", C.c_double * 3),
        ("basis_u", C.c_double * 3),
        ("basis_v", C.c_double * 3),
        ("extent", C.c_double * 2),
        ("radial", Table),
    ]


class Particles(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("dimension", C.c_int32),
        ("pad_", C.c_int32),
        ("x", C.c_void_p),
        ("dir", C.c_void_p),
        ("perp", C.c_void_p),
        ("p", C.c_void_p),
        ("alive", C.c_void_p),
        ("t", C.c_void_p),
    ]


class Draws(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("n_slots", C.c_int32),
        ("replay", C.c_void_p),
        ("n_steps", C.c_int64),
        ("seed", C.c_uint64),
        ("stream", C.c_uint64),
    ]


class Bins(C.Structure):
    _fields_ = [
        ("n_edges", C.c_int32 * 4),
        ("edges", C.c_void_p * 4),
        ("stride", C.c_int64 * 4),
        ("time_stride", C.c_int64),
        ("mode_stride", C.c_int64),
        ("direction_variable", C.c_int32),
        ("pad_", C.c_int32),
    ]


class Problem(C.Structure):
    _fields_ = [
        ("geometry", Geometry),
        ("source", Source),
        ("material", Material),
        ("bins", Bins),
        ("n_particles", C.c_int64),
        ("particle_offset", C.c_uint64),
        ("seed", C.c_uint64),
        ("n_times", C.c_int32),
        ("time_steps", C.c_int32),
        ("dt", C.c_double),
        ("precision", C.c_int32),
        ("block_size", C.c_int32),
        ("particles_per_thread", C.c_int32),
        ("hist_mode", C.c_int32),
    ]


class BatchItem(C.Structure):
    _fields_ = [
        ("material", Material),
        ("source", Source),
        ("n_particles", C.c_int64),
        ("particle_offset", C.c_uint64),
        ("seed", C.c_uint64),
        ("E", C.c_void_p),
        ("counters", C.c_void_p),
    ]


class Batch(C.Structure):
    _fields_ = [
        ("geometry", Geometry),
        ("bins", Bins),
        ("n_times", C.c_int32),
        ("time_steps", C.c_int32),
        ("dt", C.c_double),
        ("precision", C.c_int32),
        ("block_size", C.c_int32),
        ("particles_per_thread", C.c_int32),
        ("hist_mode", C.c_int32),
        ("event_batch", C.c_int32),
        ("n_items", C.c_int32),
        ("items", C.POINTER(BatchItem)),
    ]


class AdjointItem(C.Structure):
    _fields_ = [
        ("weight", C.c_void_p),
        ("grad", C.c_void_p),
        ("dhist", C.c_void_p),
        ("cv", C.c_double),
        ("dlogSigma_dlc", C.c_double),
        ("dlogsigma_dlc", C.c_void_p),
        ("n_dlogsigma", C.c_int64),
    ]


class Tuning(C.Structure):
    _fields_ = [
        ("block_size", C.c_int32),
        ("particles_per_thread", C.c_int32),
        ("hist_mode", C.c_int32),
        ("event_batch", C.c_int32),
        ("ms", C.c_double),
        ("ms_default", C.c_double),
    ]


PSDF = {"exp": 0, "power_law": 1, "gaussian": 2, "triangular": 3, "low_pass": 4, "VonKarman": 5}


class MaterialParams(C.Structure):
    _fields_ = [
        ("dimension", C.c_int32),
        ("acoustics", C.c_int32),
        ("v", C.c_double * 2),
        ("frequency", C.c_double),
        ("cv", C.c_double * 3),
        ("corr", C.c_double * 3),
        ("spectral_law", C.c_int32),
        ("pad_", C.c_int32),
        ("lc", C.c_double),
        ("nu", C.c_double),
        ("Q", C.c_double * 2),
        ("knot_threshold", C.c_double),
        ("n_theta", C.c_int64),
        ("rho", C.c_double),
        ("lds_tol", C.c_double),
        ("lds_max_knots", C.c_int32),
        ("pad2_", C.c_int32),
    ]


class MaterialHost(C.Structure):
    _fields_ = [
        ("Sigma", C.c_double * 4),
        ("Sigma_prime", C.c_double * 4),
        ("mft", C.c_double * 2),
        ("P2P", C.c_double),
        ("S2S", C.c_double),
        ("n_knots", C.c_int64 * 4),
        ("x", C.POINTER(C.c_double) * 4),
        ("y", C.POINTER(C.c_double) * 4),
        ("extrap", C.c_int32),
        ("pad_", C.c_int32),
    ]


class SubstepInfo(C.Structure):
    _fields_ = [
        ("T_eff", C.c_double),
        ("dt", C.c_double),
        ("n_sub", C.c_int64),
        ("pkill_p", C.c_double),
        ("pkill_s", C.c_double),
    ]


LIB_NAME = "librtmc.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# exported symbols of include/rtmc.h (checked by the CPU test suite)
EXPORTS = (
    "rtmc_last_error",
    "rtmc_abi_version",
    "rtmc_init_particles",
    "rtmc_propagate_smalldt",
    "rtmc_propagate_events",
    "rtmc_observe",
    "rtmc_forward",
    "rtmc_forward_noalloc",
    "rtmc_forward_batched",
    "rtmc_adjoint_batched",
    "rtmc_autotune_batched",
    "rtmc_prepare_material_host",
    "rtmc_free_material_host",
    "rtmc_prepare_material",
    "rtmc_free_material",
    "rtmc_prepare_materials",
    "rtmc_lc_derivatives",
    "rtmc_zoeppritz_air",
    "rtmc_compress_table",
    "rtmc_attach_lds_tables",
    "rtmc_probe_atomics",
    "rtmc_set_trace_timing",
    "rtmc_trace_times",
)

_lib = None


class RTMCError(RuntimeError):
    """A librtmc call returned a non-zero status."""


def _declare(lib):
    P = C.POINTER
    lib.rtmc_last_error.restype = C.c_char_p
    lib.rtmc_last_error.argtypes = []
    lib.rtmc_abi_version.restype = C.c_int
    lib.rtmc_abi_version.argtypes = []
    lib.rtmc_init_particles.restype = C.c_int
    lib.rtmc_init_particles.argtypes = [P(Source), C.c_int32, P(Draws), P(Particles), C.c_void_p]
    lib.rtmc_propagate_smalldt.restype = C.c_int
    lib.rtmc_propagate_smalldt.argtypes = [P(Material), P(Geometry), P(Particles), C.c_double,
                                           C.c_double, P(Draws), P(SubstepInfo), C.c_void_p]
    lib.rtmc_propagate_events.restype = C.c_int
    lib.rtmc_propagate_events.argtypes = [P(Material), P(Particles), C.c_double, P(Draws), C.c_void_p]
    lib.rtmc_observe.restype = C.c_int
    lib.rtmc_observe.argtypes = [P(Geometry), P(Bins), C.c_int32, P(Particles), C.c_void_p,
                                 C.c_int64, C.c_void_p]
    lib.rtmc_forward.restype = C.c_int
    lib.rtmc_forward.argtypes = [P(Problem), C.c_void_p, C.c_void_p, C.c_void_p]
    lib.rtmc_forward_noalloc.restype = C.c_int
    lib.rtmc_forward_noalloc.argtypes = [P(Problem), C.c_void_p, C.c_void_p, C.c_void_p]
    lib.rtmc_forward_batched.restype = C.c_int
    lib.rtmc_forward_batched.argtypes = [P(Batch), C.c_void_p]
    lib.rtmc_adjoint_batched.restype = C.c_int
    lib.rtmc_adjoint_batched.argtypes = [P(Batch), P(AdjointItem), C.c_void_p]
    lib.rtmc_autotune_batched.restype = C.c_int
    lib.rtmc_autotune_batched.argtypes = [P(Batch), C.c_double, P(Tuning), C.c_void_p]
    lib.rtmc_prepare_material_host.restype = C.c_int
    lib.rtmc_prepare_material_host.argtypes = [P(MaterialParams), P(MaterialHost)]
    lib.rtmc_free_material_host.restype = None
    lib.rtmc_free_material_host.argtypes = [P(MaterialHost)]
    lib.rtmc_prepare_material.restype = C.c_int
    lib.rtmc_prepare_material.argtypes = [P(MaterialParams), P(Material), P(C.c_void_p)]
    lib.rtmc_free_material.restype = C.c_int
    lib.rtmc_free_material.argtypes = [C.c_void_p]
    lib.rtmc_prepare_materials.restype = C.c_int
    lib.rtmc_prepare_materials.argtypes = [P(MaterialParams), C.c_int32, P(Material), P(C.c_void_p)]
    lib.rtmc_lc_derivatives.restype = C.c_int
    lib.rtmc_lc_derivatives.argtypes = [P(MaterialParams), C.c_int32, C.c_int64, C.c_double, P(C.c_double),
                                        P(C.c_double)]
    lib.rtmc_zoeppritz_air.restype = C.c_int
    lib.rtmc_zoeppritz_air.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int64, P(C.c_double),
                                       P(C.c_double), P(C.c_double)]
    lib.rtmc_compress_table.restype = C.c_int
    lib.rtmc_compress_table.argtypes = [P(C.c_double), P(C.c_double), C.c_int64, C.c_double, C.c_int32, P(LdsKnot),
                                        P(C.c_int32), P(C.c_double)]
    lib.rtmc_attach_lds_tables.restype = C.c_int
    lib.rtmc_attach_lds_tables.argtypes = [P(Material), C.c_int32, C.c_double, C.c_int32, P(C.c_void_p), C.c_void_p]
    lib.rtmc_set_trace_timing.restype = C.c_int
    lib.rtmc_set_trace_timing.argtypes = [C.c_int32]
    lib.rtmc_trace_times.restype = C.c_int
    lib.rtmc_trace_times.argtypes = [P(C.c_double), C.c_int32, P(C.c_int32)]
    lib.rtmc_probe_atomics.restype = C.c_int
    lib.rtmc_probe_atomics.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_int64, P(C.c_double), C.c_void_p]
    return lib


def load(path: str | None = None):
    """Load librtmc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RTMCError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = _declare(C.CDLL(p))
    v = lib.rtmc_abi_version()
    if v != ABI_VERSION:
        raise RTMCError(f"librtmc ABI version {v}, expected {ABI_VERSION}")
    if path is None:
        _lib = lib
    return lib


def check(status: int, what: str) -> None:
    if status != OK:
        msg = load().rtmc_last_error().decode(errors="replace")
        raise RTMCError(f"{what} failed with status {status}: {msg}")

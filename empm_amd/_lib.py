"""ctypes binding of libempm.so (the C-ABI declared in include/empm.h).

This module is the only place that touches the shared library.  It loads the
in-tree ``empm_amd/libempm.so`` and raises if it is missing: there is no CPU
or PyTorch fallback for the substep.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libempm.so")

EMPM_ABI_VERSION = 8
EMPM_MAX_COLLIDERS = 16

EMPM_OK = 0
EMPM_E_INVALID_ARG = -1
EMPM_E_HIP = -2
EMPM_E_OOM = -3
EMPM_E_UNSUPPORTED_MATERIAL = -4
EMPM_E_NONFINITE = -5
EMPM_E_UNSUPPORTED = -6
EMPM_STATUS_NONFINITE = 1
EMPM_STATUS_CFL = 2

EMPM_ADJ_EXACT = 0
EMPM_ADJ_DROP_GRID_MASS = 1

# Every symbol include/empm.h declares (checked by the CPU test-suite).
EXPORTED_SYMBOLS = (
    "empm_abi_version",
    "empm_last_error",
    "empm_create",
    "empm_create_batch",
    "empm_destroy",
    "empm_mu_lam",
    "empm_mu_lam_adj",
    "empm_mass_from_density",
    "empm_nn_sqdist",
    "empm_nn_sqdist_workspace",
    "empm_nn_sqdist_noalloc",
    "empm_read_status",
    "empm_set_graphs",
    "empm_graph_stats",
    "empm_set_tuning",
    "empm_set_deterministic",
    "empm_get_tuning",
    "empm_nn_sqdist_segments",
    "empm_nn_sqdist_segments_workspace",
    "empm_substep_fwd",
    "empm_substep_adj",
    "empm_restart_deformation",
    "empm_rollout_fwd",
    "empm_rollout_fwd_replay",
    "empm_rollout_adj",
    "empm_batch_rollout_fwd",
    "empm_batch_rollout_fwd_replay",
    "empm_batch_rollout_adj",
    "empm_batch_sizes",
    "empm_batch_sort_by_cell",
    "empm_sort_by_cell",
    "empm_bin_record_ints",
    "empm_profile_enable",
    "empm_profile_read",
    "empm_profile_reset",
    "empm_kernel_name",
    "empm_knn",
    "empm_lbs_weights",
    "empm_lbs_bones",
    "empm_lbs_points",
)

KERNEL_NAMES = ("k_p2g", "k_grid", "k_g2p", "k_g2p_adj", "k_grid_adj", "k_p2g_adj", "k_stress", "k_g2p_adj_scatter",
                "k_bin")


class EmpmError(RuntimeError):
    """A libempm call returned a negative status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"libempm error {code}: {message}")
        self.code = code


class CreateDesc(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("n_grid", C.c_int),
        ("max_particles", C.c_int),
        ("pad0", C.c_int),
    ]


class Model(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int),
        ("n_grid", C.c_int),
        ("material", C.c_int),
        ("pad0", C.c_int),
        ("grid_lim", C.c_float),
        ("dx", C.c_float),
        ("inv_dx", C.c_float),
        ("rpic_damping", C.c_float),
        ("grid_v_damping_scale", C.c_float),
        ("gravity", C.c_float * 3),
        ("alpha", C.c_float),
        ("hardening", C.c_float),
        ("xi", C.c_float),
        ("plastic_viscosity", C.c_float),
        ("softening", C.c_float),
        ("mu", C.c_void_p),
        ("lam", C.c_void_p),
        ("yield_stress", C.c_void_p),
        ("mass", C.c_void_p),
        ("vol", C.c_void_p),
        ("selection", C.c_void_p),
    ]


class Particles(C.Structure):
    _fields_ = [("x", C.c_void_p), ("v", C.c_void_p), ("C", C.c_void_p), ("F", C.c_void_p),
                ("Fe", C.c_void_p), ("mstate", C.c_void_p)]


class Batch(C.Structure):
    _fields_ = [
        ("n_cand", C.c_int),
        ("n_particles", C.c_int),
        ("models", C.c_void_p),
        ("mu", C.c_void_p),
        ("lam", C.c_void_p),
        ("yield_stress", C.c_void_p),
        ("mass", C.c_void_p),
        ("vol", C.c_void_p),
        ("selection", C.c_void_p),
    ]


class BatchCreateDesc(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("max_cand", C.c_int),
        ("max_particles", C.c_int),
        ("max_grid", C.c_int),
    ]


class GridRec(C.Structure):
    _fields_ = [("mv_m", C.c_void_p), ("v_out", C.c_void_p), ("bins", C.c_void_p)]


class Tuning(C.Structure):
    """empm_tuning: launch parameters of the binned kernels."""
    _fields_ = [("scatter_wg", C.c_int), ("scatter_per_cu", C.c_int), ("heavy_chunks", C.c_int),
                ("dynamic_rounds", C.c_int), ("node_grid_cap", C.c_int)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    @classmethod
    def of(cls, d):
        t = cls()
        for f, _ in cls._fields_:
            setattr(t, f, int(d[f]))
        return t


class Collider(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("surface_type", C.c_int),
        ("geo_type", C.c_int),
        ("reset", C.c_int),
        ("active", C.c_int),
        ("pad0", C.c_int),
        ("start_time", C.c_float),
        ("end_time", C.c_float),
        ("point", C.c_float * 3),
        ("normal", C.c_float * 3),
        ("size", C.c_float * 3),
        ("velocity", C.c_float * 3),
        ("rot_axis", C.c_float * 3),
        ("angular_velocity", C.c_float),
        ("radius", C.c_float),
        ("half_height", C.c_float),
        ("friction", C.c_float),
        ("mask", C.c_void_p),
    ]


_lib = None


def _declare(lib):
    P = C.POINTER
    vp = C.c_void_p
    lib.empm_abi_version.restype = C.c_int
    lib.empm_abi_version.argtypes = []
    lib.empm_last_error.restype = C.c_char_p
    lib.empm_last_error.argtypes = []
    lib.empm_create.restype = C.c_int
    lib.empm_create.argtypes = [P(vp), P(CreateDesc)]
    lib.empm_create_batch.restype = C.c_int
    lib.empm_create_batch.argtypes = [P(vp), P(BatchCreateDesc)]
    lib.empm_batch_rollout_fwd.restype = C.c_int
    lib.empm_batch_rollout_fwd.argtypes = [vp, P(Batch), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                           C.c_int, P(Particles), P(GridRec), P(Collider), P(C.c_double), vp]
    lib.empm_batch_rollout_fwd_replay.restype = C.c_int
    lib.empm_batch_rollout_fwd_replay.argtypes = [vp, P(Batch), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                                  C.c_int, P(Particles), P(GridRec), vp]
    lib.empm_batch_rollout_adj.restype = C.c_int
    lib.empm_batch_rollout_adj.argtypes = [vp, P(Batch), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                           C.c_int, P(Particles), P(GridRec), P(Particles),
                                           vp, vp, vp, vp, vp, C.c_int, vp]
    lib.empm_batch_sort_by_cell.restype = C.c_int
    lib.empm_batch_sort_by_cell.argtypes = [vp, P(Batch), vp, vp, vp]
    lib.empm_batch_sizes.restype = C.c_int
    lib.empm_batch_sizes.argtypes = [P(Batch), P(C.c_longlong), P(C.c_longlong)]
    lib.empm_destroy.restype = None
    lib.empm_destroy.argtypes = [vp]
    lib.empm_mu_lam.restype = C.c_int
    lib.empm_mu_lam.argtypes = [C.c_int, vp, vp, vp, vp, vp]
    lib.empm_mu_lam_adj.restype = C.c_int
    lib.empm_mu_lam_adj.argtypes = [C.c_int, vp, vp, vp, vp, vp, vp, vp]
    lib.empm_mass_from_density.restype = C.c_int
    lib.empm_mass_from_density.argtypes = [C.c_int, vp, vp, vp, vp]
    lib.empm_nn_sqdist.restype = C.c_int
    lib.empm_nn_sqdist.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, C.c_size_t, vp]
    lib.empm_set_graphs.restype = C.c_int
    lib.empm_set_graphs.argtypes = [vp, C.c_int]
    lib.empm_graph_stats.restype = C.c_int
    lib.empm_graph_stats.argtypes = [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    lib.empm_set_tuning.restype = C.c_int
    lib.empm_set_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.empm_set_deterministic.restype = C.c_int
    lib.empm_set_deterministic.argtypes = [vp, C.c_int]
    lib.empm_get_tuning.restype = C.c_int
    lib.empm_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    lib.empm_read_status.restype = C.c_int
    lib.empm_read_status.argtypes = [vp, vp, C.c_int, C.c_int, vp]
    lib.empm_nn_sqdist_noalloc.restype = C.c_int
    lib.empm_nn_sqdist_noalloc.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp]
    lib.empm_nn_sqdist_workspace.restype = C.c_size_t
    lib.empm_nn_sqdist_workspace.argtypes = [C.c_int, C.c_int]
    lib.empm_nn_sqdist_segments.restype = C.c_int
    lib.empm_nn_sqdist_segments.argtypes = [C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, vp, vp,
                                            C.c_size_t, vp]
    lib.empm_nn_sqdist_segments_workspace.restype = C.c_size_t
    lib.empm_nn_sqdist_segments_workspace.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    lib.empm_knn.restype = C.c_int
    lib.empm_knn.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp]
    lib.empm_lbs_weights.restype = C.c_int
    lib.empm_lbs_weights.argtypes = [C.c_int, vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp]
    lib.empm_lbs_bones.restype = C.c_int
    lib.empm_lbs_bones.argtypes = [C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, vp, vp]
    lib.empm_lbs_points.restype = C.c_int
    lib.empm_lbs_points.argtypes = [C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, C.c_int, vp, vp, vp, vp, vp]
    lib.empm_restart_deformation.restype = C.c_int
    lib.empm_restart_deformation.argtypes = [vp, P(Model), vp, vp, vp, vp]
    lib.empm_substep_fwd.restype = C.c_int
    lib.empm_substep_fwd.argtypes = [vp, P(Model), P(Collider), C.c_int, C.c_float, C.c_float,
                                     P(Particles), P(Particles), vp]
    lib.empm_substep_adj.restype = C.c_int
    lib.empm_substep_adj.argtypes = [vp, P(Model), P(Collider), C.c_int, C.c_float, C.c_float,
                                     P(Particles), P(Particles), P(Particles),
                                     vp, vp, vp, vp, vp, C.c_int, vp]
    lib.empm_rollout_fwd.restype = C.c_int
    lib.empm_rollout_fwd.argtypes = [vp, P(Model), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                     C.c_int, P(Particles), P(GridRec), P(Collider), P(C.c_double), vp]
    lib.empm_rollout_fwd_replay.restype = C.c_int
    lib.empm_rollout_fwd_replay.argtypes = [vp, P(Model), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                            C.c_int, P(Particles), P(GridRec), vp]
    lib.empm_sort_by_cell.restype = C.c_int
    lib.empm_sort_by_cell.argtypes = [vp, C.c_int, vp, C.c_float, vp, vp]
    lib.empm_bin_record_ints.restype = C.c_longlong
    lib.empm_bin_record_ints.argtypes = [C.c_int, C.c_int]
    lib.empm_profile_enable.restype = C.c_int
    lib.empm_profile_enable.argtypes = [vp, C.c_uint]
    lib.empm_profile_read.restype = C.c_int
    lib.empm_profile_read.argtypes = [vp, C.c_int, P(C.c_long), P(C.c_double)]
    lib.empm_profile_reset.restype = C.c_int
    lib.empm_profile_reset.argtypes = [vp]
    lib.empm_kernel_name.restype = C.c_char_p
    lib.empm_kernel_name.argtypes = [C.c_int]
    lib.empm_rollout_adj.restype = C.c_int
    lib.empm_rollout_adj.argtypes = [vp, P(Model), P(Collider), C.c_int, P(C.c_double), C.c_double,
                                     C.c_int, P(Particles), P(GridRec), P(Particles),
                                     vp, vp, vp, vp, vp, C.c_int, vp]


def load(path: str | None = None):
    """Load libempm.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    p = path or os.environ.get("EMPM_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise ImportError(
            f"libempm.so not found at {p}: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    _declare(lib)
    ver = lib.empm_abi_version()
    if ver != EMPM_ABI_VERSION:
        raise ImportError(f"libempm ABI version {ver} != expected {EMPM_ABI_VERSION}")
    _lib = lib
    return lib


def check(rc: int):
    if rc != EMPM_OK:
        msg = load().empm_last_error().decode(errors="replace")
        raise EmpmError(rc, msg)
    return rc

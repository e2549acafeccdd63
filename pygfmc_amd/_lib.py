"""ctypes binding of ``libgfmc_hip.so`` (the C-ABI of ``include/gfmc.h``).

The library is built in-tree by ``python -m pygfmc_amd.build``.  There is no
fallback: when the library is missing, or no ROCm GPU is visible, every
device entry point raises.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgfmc_hip.so")

GFMC_FP64 = 0
GFMC_FP32 = 1
GFMC_EUNSUP = -4
GUIDE_NONE = 0
GUIDE_JASTROW = 1
GUIDE_RBM = 2
GUIDE_MLP = 3
TAG_PROPAGATE = 0
TAG_RECONFIG = 1
TAG_INIT = 2
TAG_WARMUP_PROPAGATE = 3
TAG_WARMUP_RECONFIG = 4
TAG_OPERATOR = 16
# largest population gfmc_reconfigure_static accepts (its static scratch)
RECONFIGURE_STATIC_MAX_WALKERS = 2097152

_i32p = ctypes.POINTER(ctypes.c_int32)
_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_vp = ctypes.c_void_p


class ProblemDesc(ctypes.Structure):
    """Mirror of ``gfmc_problem_desc``."""
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_edges", ctypes.c_int32),
        ("n_hx", ctypes.c_int32), ("n_jpm", ctypes.c_int32),
        ("has_jz", ctypes.c_int32), ("uniform_jz", ctypes.c_int32),
        ("has_hz", ctypes.c_int32), ("guide", ctypes.c_int32),
        ("edge_i", _i32p), ("edge_j", _i32p),
        ("jz", _f64p), ("jpm", _f64p), ("jpmpm", _f64p), ("hz", _f64p), ("hx_half", _f64p),
        ("jas_max_deg", ctypes.c_int32),
        ("jas_site_other", _i32p), ("jas_site_J", _f64p), ("jas_h", _f64p),
        ("jas_edge_other", _i32p), ("jas_edge_J", _f64p),
        ("rbm_hidden", ctypes.c_int32),
        ("rbm_W", _f32p), ("rbm_a", _f32p), ("rbm_b", _f32p),
        ("dep_off", _i32p), ("dep_slots", _i32p), ("edge_off", _i32p), ("dep_edges", _i32p),
        ("mlp_h1", ctypes.c_int32), ("mlp_h2", ctypes.c_int32),
        ("mlp_W1", _f32p), ("mlp_b1", _f32p), ("mlp_W2", _f32p), ("mlp_b2", _f32p),
        ("mlp_w3", _f32p), ("mlp_b3", _f32p),
    ]


class ProblemInfo(ctypes.Structure):
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_words", ctypes.c_int32),
        ("n_off", ctypes.c_int32), ("m_total", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("guide", ctypes.c_int32),
        ("block_k", ctypes.c_int32), ("full_recompute", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int64),
        ("lanes", ctypes.c_int32), ("lanes_ok", ctypes.c_int32),
    ]


DIAG_PAIR = 0
DIAG_SF = 1


class DiagDesc(ctypes.Structure):
    """Mirror of ``gfmc_diag_desc``."""
    _fields_ = [
        ("n_sites", ctypes.c_int32), ("n_cols", ctypes.c_int32),
        ("kind", _i32p), ("edge_off", _i32p), ("edge_i", _i32p), ("edge_j", _i32p),
        ("edge_c", _f64p), ("a", _f64p), ("b", _f64p), ("c", _f64p),
        ("n_p", ctypes.c_int32), ("m", _i32p),
    ]


# (name, restype, argtypes) of every exported entry point
SIGNATURES = [
    ("gfmc_version", ctypes.c_char_p, []),
    ("gfmc_last_error", ctypes.c_char_p, []),
    ("gfmc_problem_create", ctypes.c_int, [ctypes.POINTER(ProblemDesc), ctypes.c_int32,
                                           ctypes.POINTER(_vp)]),
    ("gfmc_problem_destroy", ctypes.c_int, [_vp]),
    ("gfmc_problem_get_info", ctypes.c_int, [_vp, ctypes.POINTER(ProblemInfo)]),
    ("gfmc_problem_set_lanes", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("gfmc_problem_set_launch", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32]),
    ("gfmc_problem_set_method", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_double]),
    ("gfmc_init_walkers_sz", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, ctypes.c_int32,
                                            ctypes.c_uint64, _vp, _vp]),
    ("gfmc_init_walkers_random", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64,
                                                ctypes.c_uint64, _vp, _vp]),
    ("gfmc_pack", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp]),
    ("gfmc_unpack", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp]),
    ("gfmc_rates", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp]),
    ("gfmc_log_psi", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    ("gfmc_propagate", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                      ctypes.c_uint32, _vp, _vp]),
    ("gfmc_reconfigure", ctypes.c_int, [ctypes.c_int64, _vp, _vp, ctypes.c_uint64,
                                        ctypes.c_uint32, ctypes.c_uint32, _f64p, _vp, _vp,
                                        _vp, _vp]),
    ("gfmc_reconfigure_work_doubles", ctypes.c_int64, [ctypes.c_int64]),
    ("gfmc_reconfigure_static", ctypes.c_int, [ctypes.c_int64, _vp, _vp, ctypes.c_uint64,
                                               ctypes.c_uint32, ctypes.c_uint32, _f64p, _vp,
                                               _vp, _vp]),
    ("gfmc_gather", ctypes.c_int, [ctypes.c_int32, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp]),
    ("gfmc_run_steps", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                      ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                      ctypes.c_uint32, _vp, _vp, _vp]),
    ("gfmc_sample_index", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
                                         _vp, _vp, _vp]),
    ("gfmc_wave_scan", ctypes.c_int, [ctypes.c_int32, _vp, _vp, _vp]),
    ("gfmc_philox_host", ctypes.c_int, [ctypes.c_int32, _u32p, _u32p, _u32p]),
    ("gfmc_batch_create", ctypes.c_int, [ctypes.POINTER(_vp), ctypes.c_int32,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_vp)]),
    ("gfmc_batch_destroy", ctypes.c_int, [_vp]),
    ("gfmc_batch_set_lanes", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("gfmc_batch_lanes_ok", ctypes.c_int, [_vp]),
    ("gfmc_batch_get_info", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int64)]),
    ("gfmc_batch_propagate", ctypes.c_int, [_vp, _vp, _vp, _vp, ctypes.c_double, _vp,
                                            ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            _vp, _vp]),
    ("gfmc_batch_reconfigure", ctypes.c_int, [_vp, _vp, _vp, ctypes.c_uint64, ctypes.c_uint32,
                                              ctypes.c_uint32, _vp, _vp, _vp, _vp]),
    ("gfmc_batch_run_steps", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            ctypes.c_int32, ctypes.c_double, _vp,
                                            ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                                            ctypes.c_uint32, _vp, _vp, _vp]),
    ("gfmc_graph_capture_steps", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp,
                                                _vp, _vp, ctypes.c_int32, ctypes.c_double,
                                                ctypes.c_double, ctypes.c_uint64,
                                                ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                                ctypes.POINTER(_vp)]),
    ("gfmc_graph_capture_batch_steps", ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                                      ctypes.c_int32, ctypes.c_double, _vp,
                                                      ctypes.c_uint64, ctypes.c_uint32,
                                                      ctypes.c_uint32, _vp, _vp, _vp,
                                                      ctypes.POINTER(_vp)]),
    ("gfmc_graph_launch", ctypes.c_int, [_vp, _vp]),
    ("gfmc_graph_destroy", ctypes.c_int, [_vp]),
    ("gfmc_fw_state_doubles", ctypes.c_int64, [ctypes.c_int32, ctypes.c_int32]),
    ("gfmc_fw_init", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp]),
    ("gfmc_fw_update", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _vp, ctypes.c_int32,
                                      _vp, _vp]),
    ("gfmc_rbm_n_params", ctypes.c_int64, [_vp]),
    ("gfmc_rbm_set_params", ctypes.c_int, [_vp, _vp, _vp]),
    ("gfmc_rbm_set_tied", ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("gfmc_rbm_untie", ctypes.c_int, [_vp]),
    ("gfmc_rbm_theta", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    ("gfmc_rbm_vmin_work_bytes", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_rbm_vmin_grad", ctypes.c_int, [_vp, ctypes.c_int64, _vp, ctypes.c_float, _vp, _vp,
                                          _vp, _vp]),
    ("gfmc_orbit_mean", ctypes.c_int, [ctypes.c_int64, _vp, ctypes.c_int32, _vp, _vp, _vp]),
    ("gfmc_orbit_sum", ctypes.c_int, [ctypes.c_int64, _vp, ctypes.c_int32, _vp, _vp, _vp, _vp]),
    ("gfmc_orbit_axpy", ctypes.c_int, [ctypes.c_int64, _vp, ctypes.c_int32, _vp, _vp, _vp,
                                       ctypes.c_float, _vp]),
    ("gfmc_rbm_orbit_O", ctypes.c_int, [_vp, ctypes.c_int64, _vp, ctypes.c_int32, _vp, _vp, _vp,
                                        _vp, _vp]),
    ("gfmc_sr_work_doubles", ctypes.c_int64, [ctypes.c_int32]),
    ("gfmc_sr_solve", ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, _vp, _vp, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int32, _vp, _vp, _vp]),
    ("gfmc_adam_step", ctypes.c_int, [ctypes.c_int64, _vp, _vp, _vp, _vp, ctypes.c_int64,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_float, _vp]),
    ("gfmc_mlp_n_params", ctypes.c_int64, [_vp]),
    ("gfmc_mlp_counters", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("gfmc_mlp_set_hop_cap", ctypes.c_int, [_vp, ctypes.c_int32]),
    ("gfmc_mlp_unfinished", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("gfmc_mlp_work_bytes", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_mlp_set_work", ctypes.c_int, [_vp, _vp, ctypes.c_int64, ctypes.c_int64, _vp]),
    ("gfmc_mlp_set_params", ctypes.c_int, [_vp, _vp, _vp]),
    ("gfmc_mlp_vmin_grad", ctypes.c_int, [_vp, ctypes.c_int64, _vp, ctypes.c_float, _vp, _vp,
                                          _vp]),
    ("gfmc_diag_create", ctypes.c_int, [ctypes.POINTER(DiagDesc), ctypes.POINTER(_vp)]),
    ("gfmc_diag_destroy", ctypes.c_int, [_vp]),
    ("gfmc_diag_ring_bytes", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_diag_work_doubles", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_diag_ring_init", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp]),
    ("gfmc_diag_values", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp]),
    ("gfmc_diag_record", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp]),
    ("gfmc_run_steps_diag", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp,
                                           _vp, ctypes.c_int32, ctypes.c_double,
                                           ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                           ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp]),
    ("gfmc_graph_capture_steps_diag", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp,
                                                     _vp, _vp, _vp, ctypes.c_int32,
                                                     ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_uint64, ctypes.c_uint32,
                                                     ctypes.c_uint32, _vp, _vp, _vp, _vp,
                                                     _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    ("gfmc_carry_bytes", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_propagate_carry", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.c_uint32, _vp, _vp, _vp, _vp]),
    ("gfmc_run_steps_carry", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp, _vp, _vp,
                                            _vp, ctypes.c_int32, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32,
                                            ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                            _vp, _vp, _vp, _vp, _vp]),
    ("gfmc_graph_capture_steps_carry", ctypes.c_int, [_vp, ctypes.c_int64, _vp, _vp, _vp, _vp,
                                                      _vp, _vp, _vp, ctypes.c_int32,
                                                      ctypes.c_double, ctypes.c_double,
                                                      ctypes.c_uint64, ctypes.c_uint32,
                                                      ctypes.c_uint32, _vp, _vp, _vp, _vp,
                                                      _vp, _vp, _vp, _vp,
                                                      ctypes.POINTER(_vp)]),
    ("gfmc_apply_operator_work_doubles", ctypes.c_int64, [_vp, ctypes.c_int64]),
    ("gfmc_apply_operator", ctypes.c_int, [_vp, ctypes.c_int64, _vp, ctypes.c_uint64,
                                           ctypes.c_uint32, ctypes.c_uint32, _vp, _vp, _vp,
                                           _vp]),
    ("gfmc_fw_operator_work_bytes", ctypes.c_int64, [_vp, _vp, ctypes.c_int64,
                                                     ctypes.c_int32]),
    ("gfmc_forward_walk_operator", ctypes.c_int, [_vp, _vp, ctypes.c_int64, _vp, _vp,
                                                  ctypes.c_int32,
                                                  ctypes.POINTER(ctypes.c_int32),
                                                  ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_uint64, ctypes.c_uint32,
                                                  ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32, _vp, _vp, _vp, _vp]),
]

_LIB = None


def load():
    """The loaded library (raises when it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m pygfmc_amd.build` "
                "(hipcc, gfx950). pygfmc_amd has no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = lib
    return _LIB


class GfmcError(RuntimeError):
    """A C-ABI call returned an error status."""


def check(rc, what):
    if rc != 0:
        msg = load().gfmc_last_error().decode(errors="replace")
        raise GfmcError(f"{what} failed ({rc}): {msg}")


def device():
    """The GPU the engine runs on (raises when no ROCm GPU is visible)."""
    if not torch.cuda.is_available():
        raise RuntimeError(
            "pygfmc_amd runs on an AMD GPU (ROCm) and none is visible; there is no "
            "CPU path.")
    return torch.device("cuda", torch.cuda.current_device())


def stream():
    """Handle of torch's current HIP stream, as the void* the C-ABI takes."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def work_doubles(n_walkers):
    """Doubles of device scratch gfmc_reconfigure needs for n_walkers."""
    return int(load().gfmc_reconfigure_work_doubles(int(n_walkers)))


def ptr(t):
    """Device pointer of a contiguous tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_contiguous():
        raise ValueError("device buffers must be contiguous")
    return ctypes.c_void_p(t.data_ptr())

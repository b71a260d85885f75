"""ctypes binding of libvamx_hip.so (the C ABI declared in include/vamx.h).

torch is imported first so that the HIP runtime torch ships is the one the
library binds to (both carry the soname libamdhip64.so.7); device buffers are
torch tensors whose ``data_ptr()`` is passed through the C ABI together with
torch's current stream.  There is no fallback: a missing library raises.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (loads torch's HIP runtime before the library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvamx_hip.so")

c_int32 = ctypes.c_int32
c_double = ctypes.c_double
c_void_p = ctypes.c_void_p


class ParGeom(ctypes.Structure):
    _fields_ = [
        ("n", c_int32),
        ("nz", c_int32),
        ("n_angles", c_int32),
        ("angles_deg", ctypes.POINTER(c_double)),
        ("det_offset", c_double),
        ("incl_deg", c_double),
        ("mu", c_double),
        ("radius", c_double),
        ("outside_mask", ctypes.POINTER(ctypes.c_uint8)),
        ("flags", c_int32),
        ("occlusion", ctypes.POINTER(ctypes.c_uint8)),
    ]


class ParTuning(ctypes.Structure):
    _fields_ = [
        ("fp_tile_angles", c_int32),
        ("fp_tile_rays", c_int32),
        ("fp_zv", c_int32),
        ("fp_reuse", c_int32),
        ("bp_tile", c_int32),
        ("bp_zv", c_int32),
        ("fp_pair", c_int32),
        ("z_split", c_int32),
    ]


class Response(ctypes.Structure):
    _fields_ = [
        ("kind", c_int32),
        ("A", c_double),
        ("K", c_double),
        ("B", c_double),
        ("M", c_double),
        ("nu", c_double),
        ("lin_M", c_double),
        ("lin_C", c_double),
        ("n_pts", c_int32),
        ("table_f", c_void_p),
        ("table_map", c_void_p),
        ("table_dmap", c_void_p),
    ]


class BclpParams(ctypes.Structure):
    _fields_ = [
        ("p", c_double),
        ("q", c_double),
        ("eps", c_double),
        ("weight", c_double),
        ("learning_rate", c_double),
        ("glb", c_double),
        ("gub", c_double),
        ("dvol", c_double),
        ("grad_op", c_int32),
    ]


IMGSEQ_MAX_COPIES = 64


class ImgSeqParams(ctypes.Structure):
    _fields_ = [
        ("n_u", c_int32),
        ("n_v", c_int32),
        ("u1", c_int32),
        ("n_copies", c_int32),
        ("v1", c_int32 * IMGSEQ_MAX_COPIES),
        ("bit_depth", c_int32),
        ("normalize", c_int32),
        ("norm_value", c_double),
        ("intensity_scale", c_double),
        ("flip_u", c_int32),
        ("flip_v", c_int32),
    ]


class RtGeom(ctypes.Structure):
    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("xmin", c_double), ("ymin", c_double), ("zmin", c_double),
        ("dx", c_double), ("dy", c_double), ("dz", c_double),
        ("nt", c_int32), ("n_angles", c_int32), ("nzr", c_int32),
        ("angles_deg", ctypes.POINTER(c_double)),
        ("incl_deg", c_double),
        ("step", c_double),
        ("max_steps", c_int32),
        ("exit_every", c_int32),
        ("solver", c_int32),
        ("param", c_int32),
        ("ray_energy0", c_double),
        ("n_grad", c_void_p),
        ("att_kind", c_int32),
        ("att_alpha", c_double),
        ("att_radius", c_double),
        ("att_grid", c_void_p),
        ("abs_grid", c_void_p),
        ("index_kind", c_int32),
        ("index_R", c_double),
        ("index_p", c_double),
        ("index_n_sur", c_double),
    ]


# constants mirrored from include/vamx.h
PAR_NO_WARP_CLIP = 1
PAR_NO_CIRCLE_CLIP = 2
PAR_PER_LANE_GEOMETRY = 4
RESP_GLF, RESP_LINEAR, RESP_IDENTITY, RESP_TABLE = 0, 1, 2, 3
RT_SYMPLECTIC_EULER, RT_FORWARD_EULER, RT_LEAPFROG, RT_RK4 = 0, 1, 2, 3
RT_CANONICAL, RT_PHYSICAL, RT_OPTICAL = 0, 1, 2
RT_ATT_NONE, RT_ATT_CYLINDER, RT_ATT_BALL, RT_ATT_GRID = 0, 1, 2, 3
RT_INDEX_GRID, RT_INDEX_LUNEBURG = 0, 1

# every symbol declared in include/vamx.h with (restype, argtypes)
_PLAN_P = c_void_p
SIGNATURES = {
    "vamx_par_create": (c_int32, [ctypes.POINTER(ParGeom), ctypes.POINTER(c_void_p)]),
    "vamx_par_forward": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_par_backward": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_par_forward_adjoint": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_par_backward_adjoint": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_par_set_tuning": (c_int32, [_PLAN_P, ctypes.POINTER(ParTuning)]),
    "vamx_par_get_tuning": (c_int32, [_PLAN_P, ctypes.POINTER(ParTuning)]),
    "vamx_bclp_workspace_bytes": (c_int32, [_PLAN_P, ctypes.POINTER(ctypes.c_int64)]),
    "vamx_bclp_step": (c_int32, [_PLAN_P, ctypes.POINTER(Response), ctypes.POINTER(BclpParams),
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "vamx_bclp_eval": (c_int32, [_PLAN_P, ctypes.POINTER(Response), ctypes.POINTER(BclpParams),
                                 c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 c_void_p, c_void_p, c_void_p]),
    "vamx_bclp_update": (c_int32, [_PLAN_P, ctypes.POINTER(BclpParams), c_void_p, c_void_p, c_void_p]),
    "vamx_response_map": (c_int32, [ctypes.POINTER(Response), c_void_p, c_void_p, c_void_p, ctypes.c_int64,
                                    c_void_p]),
    "vamx_bclp_step_noalloc": (c_int32, [_PLAN_P, ctypes.POINTER(Response), ctypes.POINTER(BclpParams),
                                         c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p]),
    "vamx_rt_create": (c_int32, [ctypes.POINTER(RtGeom), ctypes.POINTER(c_void_p)]),
    "vamx_rt_integrate": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_rt_deposit": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p]),
    "vamx_rt_stats": (c_int32, [_PLAN_P, ctypes.POINTER(ctypes.c_int64)]),
    "vamx_rt_bclp_workspace_bytes": (c_int32, [_PLAN_P, ctypes.POINTER(ctypes.c_int64)]),
    "vamx_rt_bclp_step": (c_int32, [_PLAN_P, ctypes.POINTER(Response), ctypes.POINTER(BclpParams),
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "vamx_rt_bclp_eval": (c_int32, [_PLAN_P, ctypes.POINTER(Response), ctypes.POINTER(BclpParams),
                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "vamx_imgseq_rotate": (c_int32, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_int32, c_int32,
                                     c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int64, c_void_p]),
    "vamx_imgseq_resize": (c_int32, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_int32, c_int32,
                                     c_void_p, ctypes.c_int64, ctypes.c_int64, c_void_p]),
    "vamx_imgseq_compose": (c_int32, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.POINTER(ImgSeqParams), c_void_p, c_void_p]),
    "vamx_osmo_model_update": (c_int32, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_float, c_int32,
                                         c_void_p]),
    "vamx_scale_clamp": (c_int32, [c_void_p, ctypes.c_int64, c_void_p, ctypes.c_float, c_void_p]),
    "vamx_scale_clamp_segments": (c_int32, [c_void_p, ctypes.c_int64, c_void_p, c_int32, ctypes.c_float, c_void_p,
                                            c_void_p]),
    "vamx_ver_workspace_bytes": (c_int32, [ctypes.POINTER(ctypes.c_int64)]),
    "vamx_ver": (c_int32, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int64, c_void_p, c_void_p,
                           c_void_p]),
    "vamx_order_stats_workspace_bytes": (c_int32, [ctypes.POINTER(ctypes.c_int64)]),
    "vamx_order_stats": (c_int32, [c_void_p, ctypes.c_int64, ctypes.c_int64, c_void_p, c_void_p, c_void_p]),
    "vamx_rt_record_counts": (c_int32, [_PLAN_P, c_void_p, c_void_p]),
    "vamx_rt_record": (c_int32, [_PLAN_P, c_void_p, c_void_p, c_void_p, c_void_p]),
    "vamx_sino_resample": (c_int32, [c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p,
                                     ctypes.c_int64, c_void_p, c_void_p]),
    "vamx_csr_spmm": (c_int32, [c_void_p, c_void_p, c_void_p, ctypes.c_int64, ctypes.c_int64, c_int32, c_void_p,
                                c_void_p, c_void_p]),
    "vamx_par_forward_batched": (c_int32, [ctypes.POINTER(c_void_p), c_int32, ctypes.POINTER(c_void_p),
                                           ctypes.POINTER(c_void_p), c_void_p]),
    "vamx_par_backward_batched": (c_int32, [ctypes.POINTER(c_void_p), c_int32, ctypes.POINTER(c_void_p),
                                            ctypes.POINTER(c_void_p), c_void_p]),
    "vamx_batch_create": (c_int32, [ctypes.POINTER(c_void_p), c_int32, ctypes.POINTER(c_void_p)]),
    "vamx_batch_forward": (c_int32, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_void_p]),
    "vamx_batch_backward": (c_int32, [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p), c_void_p]),
    "vamx_batch_destroy": (None, [c_void_p]),
    "vamx_destroy": (None, [_PLAN_P]),
    "vamx_last_error": (ctypes.c_char_p, []),
    "vamx_version": (c_int32, []),
}

_lib = None
MISSING: list = []


class VamxError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load libvamx_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise VamxError(
            "libvamx_hip.so is not built (expected at %s); run `python -m vamtoolbox_amd.build`" % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name, None)
        if fn is None:
            MISSING.append(name)
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().vamx_last_error()
        raise VamxError("%s failed (status %d): %s" % (what, rc, msg.decode() if msg else ""))


def call(name: str, *args) -> None:
    lib = load()
    check(getattr(lib, name)(*args), name)


def stream_ptr(device=None) -> int:
    """HIP stream handle of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor) -> int:
    return t.data_ptr()

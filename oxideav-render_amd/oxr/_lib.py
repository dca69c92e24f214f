"""Loader of liboxr.so (the HIP path tracer). Fails loudly: there is no
CPU fallback anywhere in the product."""
import ctypes as C
import os

from . import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboxr.so")
_LIB = None


class BackendError(RuntimeError):
    """Device / runtime failure: the reference's Error::Backend
    (oxideav-render src/error.rs:38-43)."""


class InvalidOptions(ValueError):
    """Error::InvalidOptions (src/error.rs:31-36)."""


def lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise BackendError(f"liboxr.so not built at {LIB_PATH}: run __graft_entry__.build() "
                           "(make -C oxideav-render_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    u32 = C.c_uint32
    fp = C.POINTER(C.c_float)
    dp = C.POINTER(C.c_double)
    bp = C.POINTER(C.c_uint8)
    up = C.POINTER(C.c_uint32)
    sig = {
        "oxr_last_error": (C.c_char_p, []),
        "oxr_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "oxr_pt_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "oxr_pt_destroy": (C.c_int, [vp]),
        "oxr_pt_upload_scene": (C.c_int, [vp, C.POINTER(A.SceneDesc)]),
        "oxr_pt_set_camera": (C.c_int, [vp, C.POINTER(A.Camera), C.POINTER(C.c_int)]),
        "oxr_pt_set_options": (C.c_int, [vp, C.POINTER(A.PtOptions), C.POINTER(C.c_int)]),
        "oxr_pt_reset": (C.c_int, [vp]),
        "oxr_pt_set_window": (C.c_int, [vp, u32, u32, u32, u32]),
        "oxr_pt_refine": (C.c_int, [vp, u32]),
        "oxr_pt_samples": (C.c_int, [vp, up]),
        "oxr_pt_read_accum": (C.c_int, [vp, dp, dp]),
        "oxr_pt_resolve_hdr": (C.c_int, [vp, C.POINTER(A.Display), fp]),
        "oxr_pt_resolve_rgba8": (C.c_int, [vp, C.POINTER(A.Display), bp]),
        "oxr_pt_set_timing": (C.c_int, [vp, C.c_int]),
        "oxr_pt_stats_get": (C.c_int, [vp, C.POINTER(A.PtStats)]),
        "oxr_pt_autotune": (C.c_int, [vp, u32, C.POINTER(A.TuneResult)]),
        "oxr_pt_set_environment": (C.c_int, [vp, fp, u32, u32, C.c_float]),
        "oxr_pt_env_eval": (C.c_int, [vp, u32, fp, fp, fp]),
        "oxr_pt_env_sample": (C.c_int, [vp, u32, fp, fp, fp, bp]),
        "oxr_pt_tex_sample": (C.c_int, [vp, u32, C.c_int, u32, fp, fp, fp]),
        "oxr_pt_material_eval": (C.c_int, [vp, u32, up, fp, fp, fp, bp, fp]),
        "oxr_sobol_owen_4d": (C.c_int, [C.c_int, u32, up, up, fp]),
        "oxr_trace_closest": (C.c_int, [vp, u32, C.POINTER(A.Ray), u32, C.POINTER(A.Hit)]),
        "oxr_trace_any": (C.c_int, [vp, u32, C.POINTER(A.Ray), bp]),
        "oxr_trace_closest_device": (C.c_int, [vp, u32, vp, vp, vp]),
        "oxr_trace_device_check": (C.c_int, [vp]),
        "oxr_scene_prebuild": (C.c_int, [vp, u32, vp]),
        "oxr_prebuilt_destroy": (C.c_int, [vp]),
        "oxr_srv_add_prebuilt": (C.c_int, [vp, vp, vp]),
        "oxr_srv_remove_scene": (C.c_int, [vp, u32]),
        "oxr_trace_closest_device_noalloc": (C.c_int, [vp, u32, vp, vp, vp]),
        "oxr_pt_bvh_info": (C.c_int, [vp, C.POINTER(A.BvhInfo)]),
        "oxr_pt_bvh_export": (C.c_int, [vp, vp, vp]),
        "oxr_bsdf_eval": (C.c_int, [C.c_int, u32, C.POINTER(A.MaterialSample), fp, fp, bp, fp, fp, fp]),
        "oxr_bsdf_sample": (C.c_int, [C.c_int, u32, C.POINTER(A.MaterialSample), fp, fp, bp, fp, fp, fp, fp, bp]),
        "oxr_srv_create": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(vp)]),
        "oxr_srv_destroy": (C.c_int, [vp]),
        "oxr_srv_add_scene": (C.c_int, [vp, C.POINTER(A.SceneDesc), up]),
        "oxr_srv_submit": (C.c_int, [vp, u32, C.POINTER(A.Camera), C.POINTER(A.PtOptions), up]),
        "oxr_srv_step": (C.c_int, [vp, up, up]),
        "oxr_srv_job_status": (C.c_int, [vp, u32, C.POINTER(A.JobStatus)]),
        "oxr_srv_read_accum": (C.c_int, [vp, u32, dp, dp]),
        "oxr_srv_release": (C.c_int, [vp, u32]),
        "oxr_srv_stats_get": (C.c_int, [vp, C.POINTER(A.SrvStats)]),
        "oxr_guide_default_config": (C.c_int, [C.POINTER(A.GuideConfig)]),
        "oxr_guide_create": (C.c_int, [C.c_int, C.POINTER(A.GuideConfig), C.POINTER(vp)]),
        "oxr_guide_destroy": (C.c_int, [vp]),
        "oxr_guide_num_params": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "oxr_guide_get_params": (C.c_int, [vp, fp, fp]),
        "oxr_guide_set_params": (C.c_int, [vp, fp, fp]),
        "oxr_guide_infer": (C.c_int, [vp, u32, fp, fp, fp]),
        "oxr_guide_pdf": (C.c_int, [vp, u32, fp, fp, fp]),
        "oxr_guide_sample": (C.c_int, [vp, u32, fp, fp, fp, fp]),
        "oxr_guide_gradient": (C.c_int, [vp, u32, fp, fp, fp, fp, fp, dp]),
        "oxr_guide_train": (C.c_int, [vp, u32, fp, fp, fp, u32, dp]),
        "oxr_guide_stats_get": (C.c_int, [vp, C.POINTER(A.GuideStats)]),
        "oxr_guide_collected": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_float), C.POINTER(C.c_uint32)]),
        "oxr_pt_set_guide": (C.c_int, [vp, vp, C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


EXPORTED = [
    "oxr_last_error", "oxr_device_count", "oxr_pt_create", "oxr_pt_destroy", "oxr_pt_upload_scene",
    "oxr_pt_set_camera", "oxr_pt_set_options", "oxr_pt_reset", "oxr_pt_set_window", "oxr_pt_refine", "oxr_pt_samples",
    "oxr_pt_read_accum", "oxr_pt_resolve_hdr", "oxr_pt_resolve_rgba8", "oxr_pt_set_timing", "oxr_pt_stats_get",
    "oxr_pt_autotune", "oxr_pt_set_environment", "oxr_pt_env_eval", "oxr_pt_tex_sample", "oxr_pt_material_eval", "oxr_pt_env_sample",
    "oxr_sobol_owen_4d", "oxr_trace_closest", "oxr_trace_any", "oxr_trace_closest_device", "oxr_trace_device_check", "oxr_trace_closest_device_noalloc", "oxr_pt_bvh_info",
    "oxr_pt_bvh_export", "oxr_bsdf_eval", "oxr_bsdf_sample",
    "oxr_srv_create", "oxr_srv_destroy", "oxr_srv_add_scene", "oxr_srv_submit", "oxr_srv_step",
    "oxr_srv_job_status", "oxr_srv_read_accum", "oxr_srv_release", "oxr_srv_stats_get",
    "oxr_scene_prebuild", "oxr_prebuilt_destroy", "oxr_srv_add_prebuilt", "oxr_srv_remove_scene",
    "oxr_guide_default_config", "oxr_guide_create", "oxr_guide_destroy", "oxr_guide_num_params",
    "oxr_guide_get_params", "oxr_guide_set_params", "oxr_guide_infer", "oxr_guide_pdf", "oxr_guide_sample",
    "oxr_guide_gradient", "oxr_guide_train", "oxr_guide_stats_get", "oxr_guide_collected", "oxr_pt_set_guide",
]


def check(rc, what=""):
    if rc != A.OXR_OK:
        msg = lib().oxr_last_error().decode(errors="replace")
        if rc == A.OXR_ERR_INVALID:
            raise InvalidOptions(f"{what}: {msg}")
        raise BackendError(f"{what}: {msg} (status {rc})")


def device_count():
    n = C.c_int(0)
    check(lib().oxr_device_count(C.byref(n)), "oxr_device_count")
    return n.value

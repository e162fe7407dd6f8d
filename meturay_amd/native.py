"""ctypes binding of libmrh.so, the C-ABI in include/mrh.h.

The library must have been built (``python -m meturay_amd.build``); loading
fails loudly otherwise — there is no fallback path.  Device buffers are
passed as raw pointers; the helpers below accept torch CUDA tensors (torch is
only the device-memory allocator here) or integers.
"""
from __future__ import annotations

import ctypes as C
import os
import re
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libmrh.so")
HEADER = os.path.join(ROOT, "include", "mrh.h")

OK, ERR_INVALID, ERR_HIP, ERR_OOM, ERR_UNSUPPORTED, ERR_BUDGET = range(6)
BATCH_NULL, BATCH_LAMBERT, BATCH_METAL, BATCH_DIELECTRIC, BATCH_EMISSIVE = range(5)
NO_HIT = 0xFFFFFFFF

_P = C.c_void_p
_U32 = C.c_uint32
_I32 = C.c_int32
_F = C.c_float


def key(batch: int, idx: int) -> int:
    """HitKey = batch:8 | id:24 (RayLib/HitStructs.h:42-77)."""
    return ((batch & 0xFF) << 24) | (idx & 0xFFFFFF)


class MrhError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mrh error {code}: {msg}")
        self.code = code


class LbvhView(C.Structure):
    _fields_ = [("n", _U32), ("bits", _U32), ("root", _U32), ("max_depth", _U32),
                ("d_sorted_codes", _P), ("d_sorted_idx", _P), ("d_left", _P), ("d_right", _P),
                ("d_parent", _P), ("d_node_aabb", _P), ("d_nodes4", _P), ("root4", _U32),
                ("reserved", _U32)]


class CameraDesc(C.Structure):
    _fields_ = [("position", _F * 3), ("gaze", _F * 3), ("up", _F * 3), ("fov_x", _F),
                ("near_plane", _F), ("far_plane", _F)]


class RenderParams(C.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("samples", _U32), ("max_depth", _U32),
                ("rr_start", _U32), ("nee", _U32), ("seed", _U32), ("count_traversal", _U32),
                ("direct_light_mis", _U32)]


class RenderStats(C.Structure):
    _fields_ = [("segments", C.c_uint64), ("bounce_rays", _U32 * 32), ("bounces", _U32),
                ("render_ms", C.c_double), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64)]


class TuneResult(C.Structure):
    _fields_ = [("best_lds", _U32), ("best_bpc", _U32), ("best_ms", C.c_double), ("trials", _U32),
                ("trial_lds", _U32 * 32), ("trial_bpc", _U32 * 32), ("trial_ms", C.c_double * 32),
                ("best_width", _U32), ("trial_width", _U32 * 32),
                ("best_block", _U32), ("best_tile", _U32), ("trial_block", _U32 * 32), ("trial_tile", _U32 * 32),
                ("reps", _U32)]


class FilterDesc(C.Structure):
    _fields_ = [("type", _U32), ("radius", _F), ("alpha", _F), ("b", _F), ("c", _F)]


FILTER_TYPES = {"Box": 0, "Tent": 1, "Gaussian": 2, "Mitchell": 3}


def filter_desc(kind: str, radius: float, alpha: float = 0.0, b: float = 0.0, c: float = 0.0) -> FilterDesc:
    """mrh_filter_desc of a "type" / "radius" (+ "alpha" or "b", "c") filter."""
    return FilterDesc(FILTER_TYPES[kind], radius, alpha, b, c)


class WfpgFieldsDesc(C.Structure):
    _fields_ = [("field_x", _U32), ("field_y", _U32), ("bin_count", _U32), ("ray_count", _U32),
                ("filter_alpha", _F), ("fields", _P), ("bin_offsets", _P), ("ray_ids", _P), ("xi", _P),
                ("eval_uv", _P), ("out_uv", _P), ("out_dir", _P), ("out_pdf", _P), ("out_eval_pdf", _P),
                ("out_filtered", _P), ("bxdf_field", _P), ("seed", _U32)]


class SvoPayloadDesc(C.Structure):
    _fields_ = [("tris", _P), ("codes", _P), ("prims", _P), ("n_records", _U32), ("vertices", _P),
                ("n_vertices", _U32), ("node_irradiance", _P), ("node_normals", _P), ("node_normal_length", _P),
                ("vertex_mask", _P)]


class WfpgTraceDesc(C.Structure):
    _fields_ = [("field_x", _U32), ("field_y", _U32), ("bin_count", _U32), ("max_query_level_offset", _U32),
                ("boundary_radiance", _F), ("bin_origins", _P), ("bin_jitter", _P), ("node_irradiance", _P),
                ("node_normals", _P), ("out_fields", _P), ("out_nodes", _P), ("out_t", _P)]


class SvoView(C.Structure):
    _fields_ = [("depth", _U32), ("nodes", _U32), ("level_count", _U32 * 11),
                ("level_offset", _U32 * 12), ("cube", _F * 4), ("d_codes", _P),
                ("d_first_child", _P), ("d_child_mask", _P), ("d_parent", _P), ("d_sum", _P),
                ("d_count", _P), ("d_radiance", _P), ("d_node_count", _P)]


class GuideParams(C.Structure):
    _fields_ = [("iterations", _U32), ("svo_depth", _U32), ("guide_ratio", _F),
                ("min_samples", _U32)]


class WfpgParams(C.Structure):
    _fields_ = [("iterations", _U32), ("svo_depth", _U32), ("guide_ratio", _F), ("min_bin_level", _U32),
                ("bin_ray_count", _U32), ("filter_alpha", _F), ("boundary_radiance", _F), ("product", _U32)]


class WfpgStats(C.Structure):
    _fields_ = [("bins", C.c_uint64), ("binned_rays", C.c_uint64), ("guided_bounces", _U32),
                ("octree_nodes", _U32), ("voxel_records", _U32)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_uint64), ("wave_paths", _U32), ("max_slots", _U32),
                ("use_graphs", _U32), ("timeout_ms", _U32), ("policy", _U32)]


class Request(C.Structure):
    _fields_ = [("scene", _P), ("camera", CameraDesc), ("params", RenderParams), ("guide", GuideParams),
                ("d_image", _P), ("user_tag", C.c_uint64)]


class RequestDone(C.Structure):
    _fields_ = [("ticket", C.c_uint64), ("user_tag", C.c_uint64), ("status", _I32), ("reserved", _U32),
                ("segments", C.c_uint64), ("submit_s", C.c_double), ("admit_s", C.c_double),
                ("done_s", C.c_double)]


class ServerInfo(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("segments", C.c_uint64), ("budget_used", C.c_uint64),
                ("active", _U32), ("queued", _U32), ("live_paths", _U32), ("reserved", _U32),
                ("fixed_bytes", C.c_uint64), ("scene_bytes", C.c_uint64)]


class InstancingDesc(C.Structure):
    _fields_ = [("mesh_first", _P), ("n_meshes", _U32), ("xform", _P), ("inst_mesh", _P), ("inst_key", _P),
                ("n_instances", _U32)]


class SceneDesc(C.Structure):
    _fields_ = [("tris", _P), ("prim_key", _P), ("n", _U32),
                ("lambert", _P), ("n_lambert", _U32), ("metal", _P), ("n_metal", _U32),
                ("dielectric", _P), ("n_dielectric", _U32), ("emissive", _P),
                ("n_emissive", _U32), ("morton_bits", _U32), ("reflect", _P), ("n_reflect", _U32),
                ("instancing", _P), ("normals", _P)]


class SceneBuffers(C.Structure):
    _fields_ = [("tris", _P), ("prim_key", _P), ("lambert", _P), ("metal", _P),
                ("dielectric", _P), ("emissive", _P), ("n", _U32), ("n_lambert", _U32),
                ("n_metal", _U32), ("n_dielectric", _U32), ("n_emissive", _U32),
                ("camera", CameraDesc)]


_SIGS = {
    "mrh_context_create": (_I32, [_I32, C.POINTER(_P)]),
    "mrh_context_destroy": (_I32, [_P]),
    "mrh_last_error": (C.c_char_p, [_P]),
    "mrh_sync": (_I32, [_P]),
    "mrh_stream": (_I32, [_P, C.POINTER(_P)]),
    "mrh_memcpy": (_I32, [_P, _P, _P, C.c_size_t]),
    "mrh_profile_enable": (_I32, [_P, _I32]),
    "mrh_profile_reset": (_I32, [_P]),
    "mrh_profile_get": (_I32, [_P, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "mrh_morton_codes": (_I32, [_P, _P, _U32, _U32, _P]),
    "mrh_sort_pairs": (_I32, [_P, _P, _P, _P, _P, _U32, _U32, _U32, _U32, _I32]),
    "mrh_partition_by_key": (_I32, [_P, _P, _U32, _U32, _U32, _U32, _P, _P]),
    "mrh_partition_by_key_noalloc": (_I32, [_P, _P, _U32, _U32, _U32, _U32, _P, _P]),
    "mrh_lbvh_build": (_I32, [_P, _P, _U32, _U32, C.POINTER(_P)]),
    "mrh_scene_trace_closest": (_I32, [_P, _P, _P, _U32, _U32, _F, _P, _P, _P, _P, _P]),
    "mrh_lbvh_view_get": (_I32, [_P, C.POINTER(LbvhView)]),
    "mrh_lbvh_destroy": (_I32, [_P]),
    "mrh_lbvh_rebuild": (_I32, [_P, _P, _P]),
    "mrh_trace_closest": (_I32, [_P, _P, _P, _U32, _U32, _F, _P, _P, _P, _P]),
    "mrh_trace_any": (_I32, [_P, _P, _P, _U32, _U32, _F, _P]),
    "mrh_scene_create": (_I32, [_P, C.POINTER(SceneDesc), C.POINTER(_P)]),
    "mrh_scene_destroy": (_I32, [_P]),
    "mrh_scene_bvh": (_I32, [_P, C.POINTER(_P)]),
    "mrh_scene_bytes": (_I32, [_P, C.POINTER(C.c_uint64)]),
    "mrh_render": (_I32, [_P, _P, C.POINTER(CameraDesc), C.POINTER(RenderParams), _P,
                          C.POINTER(RenderStats)]),
    "mrh_render_bytes": (_I32, [C.POINTER(RenderParams), C.POINTER(C.c_uint64)]),
    "mrh_device_alloc": (_I32, [_P, C.POINTER(_P), C.c_size_t]),
    "mrh_tune": (_I32, [_P, C.c_char_p, C.c_int64]),
    "mrh_tune_get": (_I32, [_P, C.c_char_p, C.POINTER(C.c_int64)]),
    "mrh_autotune": (_I32, [_P, _P, C.POINTER(CameraDesc), C.POINTER(RenderParams), _P, _U32,
                            C.POINTER(TuneResult)]),
    "mrh_device_free": (_I32, [_P, _P]),
    "mrh_tracer_create": (_I32, [_P, _P, C.POINTER(RenderParams), C.POINTER(_P)]),
    "mrh_tracer_generate_work": (_I32, [_P, C.POINTER(CameraDesc)]),
    "mrh_tracer_render": (_I32, [_P, C.POINTER(_I32)]),
    "mrh_tracer_finalize": (_I32, [_P, _P]),
    "mrh_tracer_reset_image": (_I32, [_P]),
    "mrh_tracer_stats": (_I32, [_P, C.POINTER(RenderStats), C.POINTER(_U32)]),
    "mrh_tracer_destroy": (_I32, [_P]),
    "mrh_tracer_set_guide": (_I32, [_P, C.POINTER(GuideParams), _P]),
    "mrh_tracer_set_wfpg": (_I32, [_P, C.POINTER(WfpgParams)]),
    "mrh_tracer_set_filter": (_I32, [_P, C.POINTER(FilterDesc)]),
    "mrh_image_write": (_I32, [C.c_char_p, _U32, _U32, _U32, _P]),
    "mrh_scene_file_load": (_I32, [C.c_char_p, _U32, C.POINTER(_P), C.c_char_p, _U32]),
    "mrh_scene_file_desc": (_I32, [_P, C.POINTER(SceneDesc)]),
    "mrh_scene_file_cameras": (_I32, [_P, C.POINTER(CameraDesc), _U32, C.POINTER(_U32)]),
    "mrh_scene_file_warnings": (C.c_char_p, [_P]),
    "mrh_scene_file_destroy": (_I32, [_P]),
    "mrh_obj_mesh_counts": (_I32, [C.c_char_p, C.POINTER(_U32), _P, _U32]),
    "mrh_filter_samples": (_I32, [_P, C.POINTER(FilterDesc), _P, _P, _U32, _U32, _U32, _P, _P]),
    "mrh_filter_resolve": (_I32, [_P, _P, _P, _U32, _U32, _P]),
    "mrh_render_filtered": (_I32, [_P, _P, C.POINTER(CameraDesc), C.POINTER(RenderParams), C.POINTER(FilterDesc),
                                   _P, C.POINTER(RenderStats)]),
    "mrh_svo_create": (_I32, [_P, C.POINTER(_P)]),
    "mrh_svo_destroy": (_I32, [_P]),
    "mrh_svo_codes": (_I32, [_P, _P, _U32, _U32, _P, _U32, _P]),
    "mrh_svo_build": (_I32, [_P, _P, _P, _U32, _P, _U32]),
    "mrh_svo_splat": (_I32, [_P, _P, _P, _P, _P, _U32]),
    "mrh_svo_filter": (_I32, [_P, _P]),
    "mrh_svo_lookup": (_I32, [_P, _P, _P, _U32, _U32, _P]),
    "mrh_svo_sample": (_I32, [_P, _P, _P, _P, _P, _U32, _P, _P]),
    "mrh_svo_pdf": (_I32, [_P, _P, _P, _P, _P, _U32, _P]),
    "mrh_svo_view_get": (_I32, [_P, C.POINTER(SvoView)]),
    "mrh_svo_bin_rays": (_I32, [_P, _P, _P, _U32, _U32, _U32, _P, _P, _P, _P, _P]),
    "mrh_wfpg_sample_fields": (_I32, [_P, C.POINTER(WfpgFieldsDesc)]),
    "mrh_svo_voxelize": (_I32, [_P, _P, _U32, C.POINTER(_F * 4), _U32, _P, _P, _U32, C.POINTER(_U32)]),
    "mrh_svo_cone_trace": (_I32, [_P, _P, _P, _U32, _F, _U32, _P, _P, _P]),
    "mrh_wfpg_trace_fields": (_I32, [_P, _P, C.POINTER(WfpgTraceDesc)]),
    "mrh_svo_payload": (_I32, [_P, _P, C.POINTER(SvoPayloadDesc)]),
    "mrh_render_wfpg": (_I32, [_P, _P, C.POINTER(CameraDesc), C.POINTER(RenderParams), C.POINTER(WfpgParams),
                               _P, C.POINTER(RenderStats), C.POINTER(WfpgStats)]),
    "mrh_render_guided": (_I32, [_P, _P, C.POINTER(CameraDesc), C.POINTER(RenderParams),
                                 C.POINTER(GuideParams), _P, _P, C.POINTER(RenderStats)]),
    "mrh_server_create": (_I32, [_P, C.POINTER(ServerParams), C.POINTER(_P)]),
    "mrh_server_context": (_I32, [_P, C.POINTER(_P)]),
    "mrh_server_submit": (_I32, [_P, C.POINTER(Request), C.POINTER(C.c_uint64)]),
    "mrh_server_step": (_I32, [_P, C.POINTER(_U32), C.POINTER(_U32)]),
    "mrh_server_poll": (_I32, [_P, C.POINTER(RequestDone), _U32, C.POINTER(_U32)]),
    "mrh_server_stats": (_I32, [_P, C.POINTER(ServerInfo)]),
    "mrh_server_cancel": (_I32, [_P, C.c_uint64]),
    "mrh_server_destroy": (_I32, [_P]),
    "mrh_gen_random_tris": (_I32, [_U32, _U32, _P]),
    "mrh_gen_random_rays": (_I32, [_U32, _U32, _P]),
    "mrh_gen_cornell": (_I32, [_U32, C.POINTER(SceneBuffers)]),
    "mrh_gen_sphere_grid": (_I32, [_U32, _U32, _U32, _U32, C.POINTER(SceneBuffers)]),
}

_lib = None


def header_symbols() -> list[str]:
    """Every function the public header declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\S.*?\b(mrh_[a-z0-9_]+)\s*\(", text, flags=re.M)))


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m meturay_amd.build` "
                "(there is no non-HIP fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(x) -> int | None:
    """Device/host address of a torch tensor, numpy array or int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError(type(x))


class Context:
    """Owns an mrh_context (device, stream, scratch, kernel timers)."""

    def __init__(self, device: int = 0):
        self.h = _P()
        self._children = weakref.WeakSet()
        rc = lib().mrh_context_create(device, C.byref(self.h))
        if rc != OK:
            raise MrhError(rc, f"mrh_context_create(device={device}) failed")
        self.device = device

    def adopt(self, obj) -> None:
        """Objects made on this context are closed before it is destroyed."""
        self._children.add(obj)

    def check(self, rc: int) -> None:
        if rc != OK:
            raise MrhError(rc, lib().mrh_last_error(self.h).decode())

    def close(self) -> None:
        if self.h:
            for child in list(self._children):
                child.close()
            lib().mrh_context_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self) -> None:
        self.check(lib().mrh_sync(self.h))

    def memcpy(self, dst, src, nbytes: int) -> None:
        self.check(lib().mrh_memcpy(self.h, ptr(dst), ptr(src), nbytes))

    def tune(self, knob: str, value: int) -> None:
        self.check(lib().mrh_tune(self.h, knob.encode(), int(value)))

    def tune_get(self, knob: str) -> int:
        v = C.c_int64(0)
        self.check(lib().mrh_tune_get(self.h, knob.encode(), C.byref(v)))
        return v.value

    def profile(self, on: bool = True) -> None:
        self.check(lib().mrh_profile_enable(self.h, 1 if on else 0))

    def profile_reset(self) -> None:
        self.check(lib().mrh_profile_reset(self.h))

    def profile_get(self, kernel: str) -> tuple[float, int]:
        ms = C.c_double(0)
        n = C.c_uint64(0)
        self.check(lib().mrh_profile_get(self.h, kernel.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def filter_samples(self, filt: FilterDesc, d_values, d_coords, n: int, width: int, height: int,
                       d_weighted, d_weights) -> None:
        """mrh_filter_samples: accumulate n samples into the filter sums."""
        self.check(lib().mrh_filter_samples(self.h, C.byref(filt), ptr(d_values), ptr(d_coords), n, width, height,
                                            ptr(d_weighted), ptr(d_weights)))

    def filter_resolve(self, d_weighted, d_weights, width: int, height: int, d_image) -> None:
        self.check(lib().mrh_filter_resolve(self.h, ptr(d_weighted), ptr(d_weights), width, height, ptr(d_image)))

    def svo_voxelize(self, d_tris, n: int, cube, depth: int, d_codes=None, capacity: int = 0, d_prims=None) -> int:
        """mrh_svo_voxelize: the number of voxel records of n triangles; the
        records (and their triangles, with d_prims) go out when d_codes holds
        them all."""
        c = (_F * 4)(*[float(x) for x in cube])
        cnt = _U32(0)
        self.check(lib().mrh_svo_voxelize(self.h, ptr(d_tris), n, C.byref(c), depth, ptr(d_codes), ptr(d_prims),
                                          capacity, C.byref(cnt)))
        return int(cnt.value)

    def wfpg_sample_fields(self, d_fields, field_x: int, field_y: int, bin_count: int, d_bin_offsets, d_ray_ids,
                           ray_count: int, d_xi, d_out_uv, d_out_pdf, alpha: float = 1.0, d_eval_uv=None,
                           d_out_eval_pdf=None, d_out_dir=None, d_out_filtered=None, d_bxdf=None,
                           seed: int = 0) -> None:
        """mrh_wfpg_sample_fields: filter each bin's radiance field, build its
        distribution and draw one guided (u,v) per ray of the bin."""
        d = WfpgFieldsDesc(field_x, field_y, bin_count, ray_count, alpha, ptr(d_fields), ptr(d_bin_offsets),
                           ptr(d_ray_ids), ptr(d_xi), ptr(d_eval_uv), ptr(d_out_uv), ptr(d_out_dir), ptr(d_out_pdf),
                           ptr(d_out_eval_pdf), ptr(d_out_filtered), ptr(d_bxdf), seed)
        self.check(lib().mrh_wfpg_sample_fields(self.h, C.byref(d)))


class Lbvh:
    """Device LBVH (mrh_lbvh)."""

    def __init__(self, ctx: Context, d_tris, n: int, bits: int = 30, handle=None):
        self.ctx = ctx
        ctx.adopt(self)
        self.owned = handle is None
        if handle is None:
            self.h = _P()
            ctx.check(lib().mrh_lbvh_build(ctx.h, ptr(d_tris), n, bits, C.byref(self.h)))
        else:
            self.h = handle

    def view(self) -> LbvhView:
        v = LbvhView()
        self.ctx.check(lib().mrh_lbvh_view_get(self.h, C.byref(v)))
        return v

    def rebuild(self, d_tris) -> None:
        self.ctx.check(lib().mrh_lbvh_rebuild(self.ctx.h, self.h, ptr(d_tris)))

    def close(self) -> None:
        if getattr(self, "owned", False) and getattr(self, "h", None):
            lib().mrh_lbvh_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def scene_buffers(gen, *args):
    """Runs a generator twice (size query, fill) and returns numpy arrays."""
    import numpy as np

    b = SceneBuffers()
    rc = gen(*args, C.byref(b))
    if rc != OK:
        raise MrhError(rc, "scene generator failed")
    out = {
        "tris": np.zeros((b.n, 9), np.float32),
        "prim_key": np.zeros(b.n, np.uint32),
        "lambert": np.zeros((max(b.n_lambert, 1), 3), np.float32),
        "metal": np.zeros((max(b.n_metal, 1), 10), np.float32),
        "dielectric": np.zeros((max(b.n_dielectric, 1), 4), np.float32),
        "emissive": np.zeros((max(b.n_emissive, 1), 3), np.float32),
    }
    b.tris = out["tris"].ctypes.data
    b.prim_key = out["prim_key"].ctypes.data
    b.lambert = out["lambert"].ctypes.data
    b.metal = out["metal"].ctypes.data
    b.dielectric = out["dielectric"].ctypes.data
    b.emissive = out["emissive"].ctypes.data
    rc = gen(*args, C.byref(b))
    if rc != OK:
        raise MrhError(rc, "scene generator failed")
    out["lambert"] = out["lambert"][: b.n_lambert]
    out["metal"] = out["metal"][: b.n_metal]
    out["dielectric"] = out["dielectric"][: b.n_dielectric]
    out["emissive"] = out["emissive"][: b.n_emissive]
    cam = b.camera
    out["camera"] = {
        "position": list(cam.position), "gaze": list(cam.gaze), "up": list(cam.up),
        "fov_x": cam.fov_x, "near": cam.near_plane, "far": cam.far_plane,
    }
    return out


def gen_cornell(subdiv: int = 4):
    return scene_buffers(lib().mrh_gen_cornell, subdiv)


def gen_sphere_grid(grid: int = 64, stacks: int = 16, slices: int = 16, seed: int = 2):
    return scene_buffers(lib().mrh_gen_sphere_grid, grid, stacks, slices, seed)


def gen_instanced_sphere_grid(grid: int = 64, stacks: int = 16, slices: int = 16, seed: int = 2,
                              meshes: int = 16):
    """The sphere grid as a two-level scene: the first `meshes` displaced
    spheres of gen_sphere_grid, moved to the origin, are the meshes; every
    grid cell is an instance of mesh (cell % meshes), placed at the cell's
    centre under a rotation about y by a seeded angle, with the flat grid's
    material for that cell as its key override.  Same camera, world
    triangle count and material mix as gen_sphere_grid."""
    import math

    import numpy as np

    flat = gen_sphere_grid(grid, stacks, slices, seed)
    per = 2 * slices * (stacks - 1)
    cells = grid * grid
    meshes = min(meshes, cells)
    tris = flat["tris"][: meshes * per].reshape(meshes, per, 3, 3).copy()
    centre = [((i // grid) - (grid - 1) * 0.5, (i % grid) - (grid - 1) * 0.5) for i in range(cells)]
    for m in range(meshes):
        tris[m, :, :, 0] -= np.float32(centre[m][0])
        tris[m, :, :, 2] -= np.float32(centre[m][1])
    rng = np.random.default_rng(seed)
    xf = np.zeros((cells, 12), np.float32)
    for i in range(cells):
        a = rng.uniform(0.0, 2.0 * math.pi)
        c, s_ = math.cos(a), math.sin(a)
        xf[i] = [c, 0.0, s_, centre[i][0], 0.0, 1.0, 0.0, 0.0, -s_, 0.0, c, centre[i][1]]
    out = dict(flat)
    out["tris"] = tris.reshape(-1, 9)
    out["prim_key"] = flat["prim_key"][: meshes * per].copy()
    out["instancing"] = {
        "mesh_first": (np.arange(meshes + 1) * per).astype(np.uint32),
        "xform": xf,
        "inst_mesh": (np.arange(cells) % meshes).astype(np.uint32),
        "inst_key": flat["prim_key"][::per][:cells].astype(np.uint32),
    }
    return out


def gen_random_tris(n: int, seed: int = 1):
    import numpy as np

    t = np.zeros((n, 9), np.float32)
    rc = lib().mrh_gen_random_tris(n, seed, t.ctypes.data)
    if rc != OK:
        raise MrhError(rc, "mrh_gen_random_tris failed")
    return t


def gen_random_rays(n: int, seed: int = 1):
    import numpy as np

    r = np.zeros((7, n), np.float32)
    rc = lib().mrh_gen_random_rays(n, seed, r.ctypes.data)
    if rc != OK:
        raise MrhError(rc, "mrh_gen_random_rays failed")
    return r


def load_scene_file(path: str, substitute: bool = False, instanced: bool = False) -> dict:
    """Reads a scene file (mrh_scene_file_load) into the dict that Scene
    takes: tris, prim_key, material tables, "camera" (the first camera),
    "cameras" (all, as dicts) and "warnings"; with instanced, the meshes and
    "instancing" of a two-level scene."""
    import math

    import numpy as np

    h = _P()
    msg = C.create_string_buffer(1024)
    flags = (1 if substitute else 0) | (2 if instanced else 0)
    rc = lib().mrh_scene_file_load(os.fsencode(path), flags, C.byref(h), msg, 1024)
    if rc != OK:
        raise MrhError(rc, msg.value.decode())
    try:
        d = SceneDesc()
        lib().mrh_scene_file_desc(h, C.byref(d))

        def take(ptr_, rows, cols, dt=np.float32):
            if rows == 0:
                return np.zeros((0, cols), dt)
            a = np.ctypeslib.as_array(C.cast(ptr_, C.POINTER(C.c_float if dt == np.float32 else C.c_uint32)),
                                      (rows * cols,))
            return np.array(a, dt).reshape(rows, cols)

        out = {
            "tris": take(d.tris, d.n, 9),
            "prim_key": take(d.prim_key, d.n, 1, np.uint32).reshape(-1),
            "lambert": take(d.lambert, d.n_lambert, 3),
            "metal": take(d.metal, d.n_metal, 10),
            "dielectric": take(d.dielectric, d.n_dielectric, 4),
            "emissive": take(d.emissive, d.n_emissive, 3),
            "reflect": take(d.reflect, d.n_reflect, 4),
        }
        if d.normals:
            out["normals"] = take(d.normals, d.n, 9)
        if d.instancing:
            idesc = C.cast(d.instancing, C.POINTER(InstancingDesc)).contents
            ni = idesc.n_instances
            out["instancing"] = {
                "mesh_first": take(idesc.mesh_first, idesc.n_meshes + 1, 1, np.uint32).reshape(-1),
                "xform": take(idesc.xform, ni, 12),
                "inst_mesh": take(idesc.inst_mesh, ni, 1, np.uint32).reshape(-1),
                "inst_key": take(idesc.inst_key, ni, 1, np.uint32).reshape(-1),
            }
        n = _U32()
        lib().mrh_scene_file_cameras(h, None, 0, C.byref(n))
        cams = (CameraDesc * max(n.value, 1))()
        lib().mrh_scene_file_cameras(h, cams, n.value, C.byref(n))
        out["cameras"] = [{"position": list(c.position), "gaze": list(c.gaze), "up": list(c.up),
                           "fov_x": c.fov_x, "near": c.near_plane, "far": c.far_plane} for c in cams[:n.value]]
        out["camera"] = out["cameras"][0] if out["cameras"] else None
        out["warnings"] = lib().mrh_scene_file_warnings(h).decode()
    finally:
        lib().mrh_scene_file_destroy(h)
    return out


IMAGE_TYPES = {"exr": 0, "pfm": 1}


def write_image(path: str, rgb, kind: str | None = None) -> None:
    """Writes a host float32 image [height, width, 3] (bottom row first) as
    EXR or PFM (mrh_image_write); the kind defaults to the file extension."""
    import numpy as np

    a = np.ascontiguousarray(rgb, dtype=np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("rgb must be [height, width, 3]")
    k = (kind or os.path.splitext(path)[1].lstrip(".")).lower()
    rc = lib().mrh_image_write(os.fsencode(path), IMAGE_TYPES[k], a.shape[1], a.shape[0], a.ctypes.data)
    if rc != OK:
        raise MrhError(rc, f"mrh_image_write({path}) failed")


def camera_desc(cam: dict) -> CameraDesc:
    c = CameraDesc()
    c.position[:] = cam["position"]
    c.gaze[:] = cam["gaze"]
    c.up[:] = cam["up"]
    c.fov_x = cam["fov_x"]
    c.near_plane = cam["near"]
    c.far_plane = cam["far"]
    return c


class Scene:
    """Resident scene (mrh_scene): triangles, HitKeys, materials and LBVH."""

    def __init__(self, ctx: Context, scene: dict, morton_bits: int = 30):
        import numpy as np

        self.ctx = ctx
        ctx.adopt(self)
        self._keep = {k: np.ascontiguousarray(v) for k, v in scene.items()
                      if isinstance(v, np.ndarray)}
        d = SceneDesc()
        k = self._keep
        d.tris = k["tris"].ctypes.data
        d.prim_key = k["prim_key"].ctypes.data
        d.n = k["tris"].shape[0]
        for name in ("lambert", "metal", "dielectric", "emissive", "reflect"):
            a = k.get(name)
            setattr(d, name, a.ctypes.data if a is not None and a.size else None)
            setattr(d, "n_" + name, a.shape[0] if a is not None else 0)
        d.morton_bits = morton_bits
        nrm = scene.get("normals")
        if nrm is not None:
            k["normals"] = np.ascontiguousarray(nrm, np.float32).reshape(-1, 9)
            d.normals = k["normals"].ctypes.data
        inst = scene.get("instancing")
        self.two_level = inst is not None
        if inst is not None:
            ik = inst.get("inst_key")
            arrs = {"mesh_first": np.ascontiguousarray(inst["mesh_first"], np.uint32),
                    "xform": np.ascontiguousarray(inst["xform"], np.float32).reshape(-1, 12),
                    "inst_mesh": np.ascontiguousarray(inst["inst_mesh"], np.uint32)}
            if ik is not None:
                arrs["inst_key"] = np.ascontiguousarray(ik, np.uint32)
            self._inst_keep = arrs
            idesc = InstancingDesc(arrs["mesh_first"].ctypes.data, len(arrs["mesh_first"]) - 1,
                                   arrs["xform"].ctypes.data, arrs["inst_mesh"].ctypes.data,
                                   arrs["inst_key"].ctypes.data if ik is not None else None,
                                   arrs["xform"].shape[0])
            self._idesc = idesc
            d.instancing = C.cast(C.pointer(idesc), _P)
        self.h = _P()
        ctx.check(lib().mrh_scene_create(ctx.h, C.byref(d), C.byref(self.h)))
        self.camera = scene.get("camera")

    def trace_closest(self, d_rays, n, pitch, tmin, d_inst, d_prim, d_t, d_bary, d_stats=None):
        """Closest hits against the whole scene (mrh_scene_trace_closest)."""
        self.ctx.check(lib().mrh_scene_trace_closest(self.ctx.h, self.h, ptr(d_rays), n, pitch, tmin, ptr(d_inst),
                                                     ptr(d_prim), ptr(d_t), ptr(d_bary), ptr(d_stats)))

    def bytes(self) -> int:
        """Device bytes the scene holds (mrh_scene_bytes)."""
        b = C.c_uint64()
        self.ctx.check(lib().mrh_scene_bytes(self.h, C.byref(b)))
        return b.value

    def bvh(self) -> Lbvh:
        h = _P()
        self.ctx.check(lib().mrh_scene_bvh(self.h, C.byref(h)))
        return Lbvh(self.ctx, None, 0, handle=h)

    def render(self, d_image, width, height, samples, max_depth, rr_start, seed=0, camera=None,
               nee=0, count_traversal=False, mis=0) -> RenderStats:
        """One frame (mrh_render); nee / mis = "NextEventEstimation" / "DirectLightMIS"."""
        cam = camera_desc(camera or self.camera)
        p = RenderParams(width, height, samples, max_depth, rr_start, nee, seed,
                         1 if count_traversal else 0, mis)
        st = RenderStats()
        self.ctx.check(lib().mrh_render(self.ctx.h, self.h, C.byref(cam), C.byref(p),
                                        ptr(d_image), C.byref(st)))
        return st

    def render_filtered(self, d_image, width, height, samples, max_depth, rr_start, seed=0, filt=None,
                        camera=None, nee=0, mis=0) -> RenderStats:
        """One frame through a reconstruction filter (mrh_render_filtered)."""
        cam = camera_desc(camera or self.camera)
        p = RenderParams(width, height, samples, max_depth, rr_start, nee, seed, 0, mis)
        st = RenderStats()
        self.ctx.check(lib().mrh_render_filtered(self.ctx.h, self.h, C.byref(cam), C.byref(p), C.byref(filt),
                                                 ptr(d_image), C.byref(st)))
        return st

    def render_guided(self, d_image, width, height, samples, max_depth, rr_start, seed=0,
                      iterations=4, svo_depth=10, guide_ratio=0.5, min_samples=512, svo=None,
                      camera=None, nee=0, mis=0) -> RenderStats:
        """Guided render (mrh_render_guided); `svo` (an Svo) keeps the last guide;
        nee / mis = "NextEventEstimation" / "DirectLightMIS"."""
        cam = camera_desc(camera or self.camera)
        p = RenderParams(width, height, samples, max_depth, rr_start, nee, seed, 0, mis)
        g = GuideParams(iterations, svo_depth, guide_ratio, min_samples)
        st = RenderStats()
        self.ctx.check(lib().mrh_render_guided(self.ctx.h, self.h, C.byref(cam), C.byref(p), C.byref(g),
                                               svo.h if svo is not None else None, ptr(d_image),
                                               C.byref(st)))
        return st

    def render_wfpg(self, d_image, width, height, samples, max_depth, rr_start, seed=0, iterations=4,
                    svo_depth=10, guide_ratio=0.5, min_bin_level=5, bin_ray_count=512, filter_alpha=1.0,
                    boundary_radiance=0.0, product=0, camera=None, nee=0, mis=0):
        """Guided render with the field-traced guide of the voxel octree
        (mrh_render_wfpg); returns (RenderStats, WfpgStats)."""
        cam = camera_desc(camera or self.camera)
        p = RenderParams(width, height, samples, max_depth, rr_start, nee, seed, 0, mis)
        w = WfpgParams(iterations, svo_depth, guide_ratio, min_bin_level, bin_ray_count, filter_alpha,
                       boundary_radiance, product)
        st, ws = RenderStats(), WfpgStats()
        self.ctx.check(lib().mrh_render_wfpg(self.ctx.h, self.h, C.byref(cam), C.byref(p), C.byref(w),
                                             ptr(d_image), C.byref(st), C.byref(ws)))
        return st, ws

    def autotune(self, d_image, width, height, samples, max_depth, rr_start, seed=0, reps=3,
                 camera=None) -> TuneResult:
        cam = camera_desc(camera or self.camera)
        p = RenderParams(width, height, samples, max_depth, rr_start, 0, seed, 0)
        res = TuneResult()
        self.ctx.check(lib().mrh_autotune(self.ctx.h, self.h, C.byref(cam), C.byref(p), ptr(d_image),
                                          reps, C.byref(res)))
        return res

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().mrh_scene_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Svo:
    """Guide octree (mrh_svo) on the context's device."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        ctx.adopt(self)
        self.h = _P()
        ctx.check(lib().mrh_svo_create(ctx.h, C.byref(self.h)))

    def codes(self, d_points, n: int, pitch: int, cube, depth: int, d_codes) -> None:
        c = (_F * 4)(*[float(x) for x in cube])
        self.ctx.check(lib().mrh_svo_codes(self.ctx.h, ptr(d_points), n, pitch, c, depth, ptr(d_codes)))

    def build(self, d_codes, n: int, cube, depth: int) -> None:
        c = (_F * 4)(*[float(x) for x in cube])
        self.ctx.check(lib().mrh_svo_build(self.ctx.h, self.h, ptr(d_codes), n, c, depth))

    def splat(self, d_codes, d_bins, d_values, n: int) -> None:
        self.ctx.check(lib().mrh_svo_splat(self.ctx.h, self.h, ptr(d_codes), ptr(d_bins), ptr(d_values), n))

    def filter(self) -> None:
        self.ctx.check(lib().mrh_svo_filter(self.ctx.h, self.h))

    def lookup(self, d_codes, n: int, d_nodes, min_count: int = 0) -> None:
        self.ctx.check(lib().mrh_svo_lookup(self.ctx.h, self.h, ptr(d_codes), n, min_count, ptr(d_nodes)))

    def bin_rays(self, d_codes, n: int, min_level: int, bin_ray_count: int, d_bins, d_order=None,
                 d_part_offsets=None, d_part_bins=None, d_part_count=None) -> None:
        """mrh_svo_bin_rays: each ray's bin node and, with d_order, the rays partitioned by bin."""
        self.ctx.check(lib().mrh_svo_bin_rays(self.ctx.h, self.h, ptr(d_codes), n, min_level, bin_ray_count,
                                              ptr(d_bins), ptr(d_order), ptr(d_part_offsets), ptr(d_part_bins),
                                              ptr(d_part_count)))

    def cone_trace(self, d_rays, n: int, aperture: float, max_query_level_offset: int, d_nodes, d_t,
                   d_is_leaf=None) -> None:
        """mrh_svo_cone_trace over rays [n][7] (origin, direction, tmin)."""
        self.ctx.check(lib().mrh_svo_cone_trace(self.ctx.h, self.h, ptr(d_rays), n, aperture, max_query_level_offset,
                                                ptr(d_nodes), ptr(d_t), ptr(d_is_leaf)))

    def payload(self, d_tris, d_codes, d_prims, n_records: int, d_vertices, n_vertices: int, d_irradiance,
                d_normals, d_lengths=None, d_mask=None) -> None:
        """mrh_svo_payload: normals and front/back radiance of every node;
        vertices whose d_mask entry is 0xFFFFFFFF are skipped."""
        d = SvoPayloadDesc(ptr(d_tris), ptr(d_codes), ptr(d_prims), n_records, ptr(d_vertices), n_vertices,
                           ptr(d_irradiance), ptr(d_normals), ptr(d_lengths), ptr(d_mask))
        self.ctx.check(lib().mrh_svo_payload(self.ctx.h, self.h, C.byref(d)))

    def trace_fields(self, field_x: int, field_y: int, bin_count: int, d_origins, d_jitter, d_irradiance, d_normals,
                     d_fields, boundary: float = 0.0, offset: int = 0, d_nodes=None, d_t=None) -> None:
        """mrh_wfpg_trace_fields: cone-traced radiance fields of the bins."""
        d = WfpgTraceDesc(field_x, field_y, bin_count, offset, boundary, ptr(d_origins), ptr(d_jitter),
                          ptr(d_irradiance), ptr(d_normals), ptr(d_fields), ptr(d_nodes), ptr(d_t))
        self.ctx.check(lib().mrh_wfpg_trace_fields(self.ctx.h, self.h, C.byref(d)))

    def sample(self, d_nodes, d_normals, d_xi, n: int, d_dirs, d_pdf) -> None:
        self.ctx.check(lib().mrh_svo_sample(self.ctx.h, self.h, ptr(d_nodes), ptr(d_normals), ptr(d_xi), n,
                                            ptr(d_dirs), ptr(d_pdf)))

    def pdf(self, d_nodes, d_normals, d_dirs, n: int, d_pdf) -> None:
        self.ctx.check(lib().mrh_svo_pdf(self.ctx.h, self.h, ptr(d_nodes), ptr(d_normals), ptr(d_dirs), n,
                                         ptr(d_pdf)))

    def view(self) -> SvoView:
        v = SvoView()
        self.ctx.check(lib().mrh_svo_view_get(self.h, C.byref(v)))
        return v

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().mrh_svo_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BorrowedContext(Context):
    """A context owned by someone else (the server's private context)."""

    def __init__(self, handle):
        self.h = handle
        self._children = weakref.WeakSet()
        self.device = None

    def close(self) -> None:
        self.h = _P()


class Server:
    """Serving engine (mrh_server): queued render requests traced together."""

    def __init__(self, ctx: Context, hbm_budget: int = 200 << 30, wave_paths: int = 1 << 24,
                 max_slots: int = 64, use_graphs: bool = True, timeout_ms: int = 0, policy: int = 0):
        self.ctx = ctx
        ctx.adopt(self)
        self.h = _P()
        p = ServerParams(hbm_budget, wave_paths, max_slots, 1 if use_graphs else 0, timeout_ms, policy)
        ctx.check(lib().mrh_server_create(ctx.h, C.byref(p), C.byref(self.h)))
        inner = _P()
        ctx.check(lib().mrh_server_context(self.h, C.byref(inner)))
        self.inner = _BorrowedContext(inner)

    def submit(self, scene: Scene, d_image, width, height, samples, max_depth, rr_start, seed=0,
               camera=None, iterations=1, svo_depth=10, guide_ratio=0.5, min_samples=512, tag=0,
               nee=0, mis=0) -> int:
        r = Request()
        r.scene = scene.h
        r.camera = camera_desc(camera or scene.camera)
        r.params = RenderParams(width, height, samples, max_depth, rr_start, nee, seed, 0, mis)
        r.guide = GuideParams(iterations, svo_depth, guide_ratio, min_samples)
        r.d_image = ptr(d_image)
        r.user_tag = tag
        t = C.c_uint64()
        self.ctx.check(lib().mrh_server_submit(self.h, C.byref(r), C.byref(t)))
        return t.value

    def cancel(self, ticket: int) -> None:
        self.ctx.check(lib().mrh_server_cancel(self.h, ticket))

    def step(self) -> tuple[int, int]:
        a, q = _U32(), _U32()
        self.ctx.check(lib().mrh_server_step(self.h, C.byref(a), C.byref(q)))
        return a.value, q.value

    def poll(self, max_items: int = 256) -> list[dict]:
        buf = (RequestDone * max_items)()
        n = _U32()
        self.ctx.check(lib().mrh_server_poll(self.h, buf, max_items, C.byref(n)))
        return [{"ticket": d.ticket, "tag": d.user_tag, "status": d.status, "segments": d.segments,
                 "submit_s": d.submit_s, "admit_s": d.admit_s, "done_s": d.done_s} for d in buf[:n.value]]

    def stats(self) -> ServerInfo:
        st = ServerInfo()
        self.ctx.check(lib().mrh_server_stats(self.h, C.byref(st)))
        return st

    def run(self) -> list[dict]:
        """Steps until nothing is active or queued; returns the completions."""
        done = []
        while True:
            a, q = self.step()
            done += self.poll()
            if a == 0 and q == 0:
                return done + self.poll()

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().mrh_server_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes bindings of the CPU oracle (liboracle.so) and of the reference's
host-compiled helpers (_ref/libmrhref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by bench.py's cpu_baseline leg, always as the checker or the baseline,
never as the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmrhref.so")

INVALID = 0xFFFFFFFF
BATCH_NULL, BATCH_LAMBERT, BATCH_METAL, BATCH_DIELECTRIC, BATCH_EMISSIVE = 0, 1, 2, 3, 4

_P = C.c_void_p
_U32 = C.c_uint32
_F = C.c_float


def build(reference: bool = False) -> None:
    """Compile liboracle.so (and _ref/libmrhref.so when asked and the
    reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)
    if reference and os.path.isdir("/root/reference/Source"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_P)


class OmCamera(C.Structure):
    _fields_ = [("position", _F * 3), ("bottom_left", _F * 3), ("right", _F * 3),
                ("up", _F * 3), ("plane_size", _F * 2), ("near_far", _F * 2)]


class OmScene(C.Structure):
    _fields_ = [("tris", _P), ("prim_key", _P), ("n", _U32),
                ("lambert", _P), ("metal", _P), ("dielectric", _P), ("emissive", _P),
                ("sorted_idx", _P), ("left", _P), ("right", _P), ("node_aabb", _P),
                ("root", _U32), ("nodes4", _P), ("root4", _U32), ("lights", _P), ("n_lights", _U32),
                ("reflect", _P), ("two_level", _P), ("tree_width", _U32), ("light_inst", _P),
                ("tbn", _P)]


class OmTwoLevel(C.Structure):
    _fields_ = [("n_inst", _U32), ("w2l", _P), ("l2w", _P), ("inst_mesh", _P), ("inst_key", _P),
                ("top_sorted_idx", _P), ("top_nodes4", _P), ("top_root4", _U32), ("n_meshes", _U32),
                ("mesh_first", _P), ("mesh_sorted_idx", _P), ("mesh_nodes4", _P), ("mesh_left", _P),
                ("mesh_right", _P), ("mesh_node_aabb", _P), ("mesh_root", _P), ("mesh_root4", _P),
                ("mesh_ibase", _P), ("mesh_abase", _P)]


class OmFilterDesc(C.Structure):
    _fields_ = [("type", _U32), ("radius", _F), ("alpha", _F), ("b", _F), ("c", _F)]


FILTERS = {"Box": 0, "Tent": 1, "Gaussian": 2, "Mitchell": 3}


def filter_desc(kind: str, radius: float, alpha: float = 0.0, b: float = 0.0, c: float = 0.0) -> OmFilterDesc:
    return OmFilterDesc(FILTERS[kind], radius, alpha, b, c)


class OmRenderParams(C.Structure):
    _fields_ = [("width", _U32), ("height", _U32), ("samples", _U32), ("max_depth", _U32),
                ("rr_start", _U32), ("seed", _U32), ("row_step", _U32), ("row_offset", _U32),
                ("threads", _U32), ("nee", _U32), ("mis", _U32)]


class OmSvo(C.Structure):
    _fields_ = [("depth", _U32), ("nodes", _U32), ("level_count", _U32 * 11),
                ("level_offset", _U32 * 12), ("cube", _F * 4),
                ("codes", _P), ("first_child", _P), ("child_mask", _P), ("parent", _P),
                ("acc", _P), ("count", _P), ("sum", _P), ("radiance", _P), ("node_count", _P),
                ("centers", _F * 192)]


class OmGuideParams(C.Structure):
    _fields_ = [("iterations", _U32), ("svo_depth", _U32), ("guide_ratio", _F), ("min_samples", _U32)]


class OmWfpgDesc(C.Structure):
    _fields_ = [("field_x", _U32), ("field_y", _U32), ("bin_count", _U32), ("ray_count", _U32),
                ("filter_alpha", _F), ("fields", _P), ("bin_offsets", _P), ("ray_ids", _P), ("xi", _P),
                ("eval_uv", _P), ("out_uv", _P), ("out_dir", _P), ("out_pdf", _P), ("out_eval_pdf", _P),
                ("out_filtered", _P), ("bxdf_field", _P), ("seed", _U32)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.om_morton_expand10.restype = _U32
        L.om_morton_expand10.argtypes = [_U32]
        L.om_morton_compose30.restype = _U32
        L.om_morton_compose30.argtypes = [_U32, _U32, _U32]
        L.om_morton_compose63.restype = C.c_uint64
        L.om_morton_compose63.argtypes = [_U32, _U32, _U32]
        L.om_morton_delta.restype = _F
        L.om_morton_delta.argtypes = [_P, _U32]
        L.om_quantise.argtypes = [_P, _P, _F, _U32, _P]
        L.om_tri_aabbs.argtypes = [_P, _U32, _P]
        L.om_extent.argtypes = [_P, _U32, _P]
        L.om_morton_codes.argtypes = [_P, _U32, _U32, _P]
        L.om_sort_pairs_u32.argtypes = [_P, _U32, _P, _P]
        L.om_sort_pairs_u64.argtypes = [_P, _U32, _P, _P]
        L.om_partition_desc.argtypes = [_P, _U32, _U32, _U32, _P, _P]
        L.om_lbvh_build.restype = C.c_int
        L.om_lbvh_build.argtypes = [_P, _U32, _U32, _P, _P, _P, _P, _P, _P, _P]
        L.om_intersect_triangle.restype = C.c_int
        L.om_intersect_triangle.argtypes = [_P, _P, _P, _P, _P]
        L.om_intersect_aabb.restype = C.c_int
        L.om_intersect_aabb.argtypes = [_P, _P, _P, _F, _F, _P]
        L.om_slab_box_test.restype = C.c_int
        L.om_slab_box_test.argtypes = [_P, _P, _P, _F, _F, _P]
        L.om_orthogonal_vector.argtypes = [_P, _P]
        L.om_space_quat.argtypes = [_P, _P, _P, _P]
        L.om_qslerp.argtypes = [_P, _P, _F, _P]
        L.om_bary_slerp.argtypes = [_P, _P, _P, _F, _F, _P]
        L.om_corner_frames.argtypes = [_P, _U32, _P]
        L.om_smooth_frame.argtypes = [_P, _P, _F, _F, _P, _P, _P]
        L.om_wfpg_gauss_weights.argtypes = [_F, _P]
        L.om_octa_wrap_int.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P]
        L.om_wfpg_filter.argtypes = [_P, _U32, _U32, _P, _P]
        L.om_wfpg_pwc.argtypes = [_P, _U32, _U32, _P, _P, _P, _P]
        L.om_wfpg_pwc_sample.argtypes = [_P, _P, _U32, _U32, _F, _F, _P]
        L.om_wfpg_pwc_pdf.restype = _F
        L.om_wfpg_pwc_pdf.argtypes = [_P, _P, _U32, _U32, _F, _F]
        L.om_wfpg_sample_fields.argtypes = [C.POINTER(OmWfpgDesc)]
        L.om_voxelize.restype = C.c_uint64
        L.om_svo_cone_trace_batch.argtypes = [C.POINTER(OmSvo), _P, _U32, _F, _U32, _P, _P, _P]
        L.om_svo_payload.argtypes = [C.POINTER(OmSvo), _P, _P, _P, _U32, _P, _U32, _P, _P]
        L.om_wfpg_trace_fields.argtypes = [C.POINTER(OmSvo), _U32, _U32, _U32, _P, _P, _P, _P, _F, _U32, _P, _P,
                                           _P]
        L.om_voxelize.argtypes = [_P, _U32, _P, _U32, _P, _P, C.c_uint64]
        L.om_trace_closest.argtypes = [_P, _U32, _P, _P, _P, _P, _U32, _P, _U32, _U32, _F,
                                       _P, _P, _P, _P]
        L.om_trace_any.argtypes = [_P, _U32, _P, _P, _P, _P, _U32, _P, _U32, _U32, _F, _P]
        L.om_trace_brute.argtypes = [_P, _U32, _P, _U32, _U32, _F, _P, _P, _P]
        L.om_affine_inverse.restype = C.c_int
        L.om_affine_inverse.argtypes = [_P, _P]
        L.om_xf_ray.argtypes = [_P, _P, _P, _P, _P]
        L.om_xf_point.argtypes = [_P, _P, _P]
        L.om_world_box.argtypes = [_P, _P, _P]
        L.om_trace2l_closest.argtypes = [_P, _P, _U32, _P, _U32, _U32, _F, _P, _P, _P, _P, _P]
        L.om_bvh4_build.restype = _U32
        L.om_bvh4_build.argtypes = [_U32, _P, _P, _P, _U32, _P]
        L.om_trace4_closest.argtypes = [_P, _U32, _P, _P, _U32, _P, _U32, _U32, _F, _P, _P, _P, _P]
        L.om_trace4_any.argtypes = [_P, _U32, _P, _P, _U32, _P, _U32, _U32, _F, _P]
        L.om_sincos_2pi_f.argtypes = [_F, _P, _P]
        L.om_hemi_uniform.argtypes = [_F, _F, _P, _P]
        L.om_reflect.argtypes = [_P, _P, _P]
        L.om_refract.restype = C.c_int
        L.om_refract.argtypes = [_P, _P, _F, _F, _P]
        L.om_triangle_normal.argtypes = [_P, _P]
        L.om_aabb_centroid.argtypes = [_P, _P]
        L.om_nudge.argtypes = [_P, _P, _P]
        L.om_rot_between_z.argtypes = [_P, _P]
        L.om_quat_rotate.argtypes = [_P, _P, _P]
        L.om_fresnel_dielectric.restype = _F
        L.om_fresnel_dielectric.argtypes = [_F, _F, _F]
        L.om_fresnel_conductor.argtypes = [_F, _P, _P, _P]
        L.om_dggx.restype = _F
        L.om_dggx.argtypes = [_F, _F]
        L.om_gsmith_single.restype = _F
        L.om_gsmith_single.argtypes = [_P, _F]
        L.om_gsmith_correlated.restype = _F
        L.om_gsmith_correlated.argtypes = [_P, _P, _F]
        L.om_vndf_pdf.restype = _F
        L.om_vndf_pdf.argtypes = [_P, _P, _F]
        L.om_vndf_sample.argtypes = [_P, _F, _F, _F, _P, _P]
        L.om_camera_setup.argtypes = [_P, _P, _P, _F, _F, _F, _U32, _U32, C.POINTER(OmCamera)]
        L.om_render.restype = C.c_uint64
        L.om_render.argtypes = [C.POINTER(OmScene), C.POINTER(OmCamera),
                                C.POINTER(OmRenderParams), _P]
        L.om_camera_ray.argtypes = [C.POINTER(OmCamera), _U32, _U32, _U32, _U32, _U32,
                                    _P, _P, C.POINTER(C.c_uint64)]
        L.om_camera_ray_at.argtypes = [C.POINTER(OmCamera), _U32, _U32, _U32, _U32, _F, _F,
                                       _P, _P]
        L.om_pcg32_stream.argtypes = [C.c_uint64, C.c_uint64, _U32, _P]
        L.om_splitmix64.restype = C.c_uint64
        L.om_splitmix64.argtypes = [C.c_uint64]
        L.om_path_uniforms.argtypes = [_U32, _U32, _U32, _P]
        L.om_svo_build.restype = C.c_int
        L.om_svo_build.argtypes = [C.POINTER(OmSvo), _P, _U32, _P, _U32]
        L.om_svo_free.argtypes = [C.POINTER(OmSvo)]
        L.om_svo_lookup.restype = _U32
        L.om_svo_lookup.argtypes = [C.POINTER(OmSvo), _U32]
        L.om_svo_lookup_guide.restype = _U32
        L.om_svo_lookup_guide.argtypes = [C.POINTER(OmSvo), _U32, _U32]
        L.om_svo_splat.argtypes = [C.POINTER(OmSvo), _P, _P, _P, _U32]
        L.om_svo_filter.argtypes = [C.POINTER(OmSvo)]
        L.om_svo_bin_rays.argtypes = [C.POINTER(OmSvo), _P, _U32, _U32, _U32, _P, _P, _P, _P, _P]
        L.om_svo_bin_rays.restype = None
        L.om_svo_codes.argtypes = [_P, _U32, _P, _U32, _P]
        L.om_dir_to_bins.argtypes = [_P, _U32, _P]
        L.om_octa_to_dirs.argtypes = [_P, _U32, _P]
        L.om_dir_to_octas.argtypes = [_P, _U32, _P]
        L.om_atan_unit.restype = _F
        L.om_atan_unit.argtypes = [_F]
        L.om_guide_sample.argtypes = [C.POINTER(OmSvo), _P, _P, _P, _U32, _P, _P]
        L.om_guide_pdf.argtypes = [C.POINTER(OmSvo), _P, _P, _P, _U32, _P]
        L.om_sample_light.restype = _U32
        L.om_sample_light.argtypes = [C.POINTER(OmScene), _P, _P, _P, C.POINTER(_F), C.POINTER(_F)]
        L.om_light_pdf_hit.restype = _F
        L.om_light_pdf_hit.argtypes = [_P, _F, _P]
        L.om_power_heuristic.restype = _F
        L.om_power_heuristic.argtypes = [_F, _F]
        L.om_filter_pixel_wh.restype = C.c_int
        L.om_filter_pixel_wh.argtypes = [_F]
        L.om_filter_weight.restype = _F
        L.om_filter_weight.argtypes = [C.POINTER(OmFilterDesc), _F, _F]
        L.om_filter_samples.argtypes = [C.POINTER(OmFilterDesc), _P, _P, _U32, _U32, _U32, _P, _P]
        L.om_filter_samples_grouped.argtypes = [C.POINTER(OmFilterDesc), _P, _P, _U32, _U32, _U32, _P, _P]
        L.om_render_samples.restype = C.c_uint64
        L.om_render_samples.argtypes = [C.POINTER(OmScene), C.POINTER(OmCamera), C.POINTER(OmRenderParams),
                                        _U32, _P, _P]
        L.om_render_guided.restype = C.c_uint64
        L.om_render_guided.argtypes = [C.POINTER(OmScene), C.POINTER(OmCamera),
                                       C.POINTER(OmRenderParams), C.POINTER(OmGuideParams), _P,
                                       C.POINTER(OmSvo)]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own host-compiled RayLib helpers (see ref_wrap.cpp)."""
    global _ref
    if _ref is None:
        R = C.CDLL(REF_PATH)
        R.ref_compose30.restype = _U32
        R.ref_compose30.argtypes = [_U32, _U32, _U32]
        R.ref_compose63.restype = C.c_uint64
        R.ref_compose63.argtypes = [_U32, _U32, _U32]
        R.ref_decompose30.argtypes = [_U32, _P]
        R.ref_decompose63.argtypes = [C.c_uint64, _P]
        R.ref_intersect_triangle.restype = C.c_int
        R.ref_intersect_triangle.argtypes = [_P, _P, _P, _P, _P]
        R.ref_intersect_aabb.restype = C.c_int
        R.ref_intersect_aabb.argtypes = [_P, _P, _P, _F, _F, _P]
        R.ref_nudge.argtypes = [_P, _P, _P, _P]
        R.ref_reflect.argtypes = [_P, _P, _P]
        R.ref_hemi_uniform.argtypes = [_F, _F, _P, _P]
        R.ref_hemi_cosine.argtypes = [_F, _F, _P, _P]
        R.ref_rot_between_z.argtypes = [_P, _P]
        R.ref_quat_rotate.argtypes = [_P, _P, _P]
        R.ref_quat_mul.argtypes = [_P, _P, _P]
        R.ref_triangle_normal.argtypes = [_P, _P]
        R.ref_triangle_aabb.argtypes = [_P, _P]
        R.ref_aabb_centroid.argtypes = [_P, _P]
        if hasattr(R, "ref_ray_transform"):
            R.ref_ray_transform.argtypes = [_P, _P, _P, _P, _P]
        R.ref_orthogonal_vector.argtypes = [_P, _P]
        R.ref_space_quat.argtypes = [_P, _P, _P, _P]
        R.ref_slerp.argtypes = [_P, _P, _F, _P]
        R.ref_bary_slerp.argtypes = [_P, _P, _P, _F, _F, _P]
        R.ref_octa_to_dir.argtypes = [_F, _F, _P]
        R.ref_dir_to_octa.argtypes = [_P, _P]
        R.ref_octa_wrap_int.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P]
        _ref = R
    return _ref


# ---------------------------------------------------------------- wrappers
def morton_codes(tris: np.ndarray, bits: int) -> np.ndarray:
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    n = tris.shape[0]
    out = np.zeros(n, dtype=np.uint32 if bits == 30 else np.uint64)
    lib().om_morton_codes(_ptr(tris), n, bits, _ptr(out))
    return out


def tri_aabbs(tris: np.ndarray) -> np.ndarray:
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    out = np.zeros((tris.shape[0], 6), dtype=np.float32)
    lib().om_tri_aabbs(_ptr(tris), tris.shape[0], _ptr(out))
    return out


def sort_pairs(keys: np.ndarray):
    keys = np.ascontiguousarray(keys)
    n = keys.shape[0]
    sk = np.zeros_like(keys)
    si = np.zeros(n, dtype=np.uint32)
    if keys.dtype == np.uint32:
        lib().om_sort_pairs_u32(_ptr(keys), n, _ptr(sk), _ptr(si))
    else:
        lib().om_sort_pairs_u64(_ptr(keys), n, _ptr(sk), _ptr(si))
    return sk, si


def partition_desc(keys: np.ndarray, begin_bit: int, end_bit: int):
    keys = np.ascontiguousarray(keys, dtype=np.uint32)
    n = keys.shape[0]
    order = np.zeros(n, dtype=np.uint32)
    offsets = np.zeros((1 << (end_bit - begin_bit)) + 1, dtype=np.uint32)
    lib().om_partition_desc(_ptr(keys), n, begin_bit, end_bit, _ptr(order), _ptr(offsets))
    return order, offsets


class LBVH:
    """Oracle LBVH in the reference's node numbering (see oracle.h)."""

    def __init__(self, tris: np.ndarray, bits: int = 30):
        self.tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        n = self.n = self.tris.shape[0]
        self.bits = bits
        self.sorted_codes = np.zeros(n, dtype=np.uint32 if bits == 30 else np.uint64)
        self.sorted_idx = np.zeros(n, dtype=np.uint32)
        ni = max(n - 1, 1)
        self.left = np.full(ni, INVALID, dtype=np.uint32)
        self.right = np.full(ni, INVALID, dtype=np.uint32)
        self.parent = np.zeros(2 * n - 1, dtype=np.uint32)
        self.node_aabb = np.zeros((2 * n - 1, 6), dtype=np.float32)
        root = C.c_uint32(0)
        rc = lib().om_lbvh_build(_ptr(self.tris), n, bits, _ptr(self.sorted_codes),
                                 _ptr(self.sorted_idx), _ptr(self.left), _ptr(self.right),
                                 _ptr(self.parent), _ptr(self.node_aabb), C.byref(root))
        if rc != 0:
            raise RuntimeError(f"om_lbvh_build failed ({rc})")
        self.root = root.value
        self._nodes4 = None
        self.root4 = None

    def nodes4(self) -> np.ndarray:
        """The 4-wide traversal tree collapsed from this LBVH (float32 [n-1, 32])."""
        if self._nodes4 is None:
            self._nodes4 = np.zeros((max(self.n - 1, 1), 32), dtype=np.float32)
            self.root4 = int(lib().om_bvh4_build(self.n, _ptr(self.left), _ptr(self.right),
                                                 _ptr(self.node_aabb), self.root, _ptr(self._nodes4)))
        return self._nodes4

    def trace(self, rays: np.ndarray, tmin: float = 0.0, stats: bool = False, width: int = 4):
        """rays: float32 [7, m] planes (ox oy oz dx dy dz tmax); width 2 walks
        the binary LBVH, 4 the collapsed 4-wide tree."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        m = rays.shape[1]
        prim = np.zeros(m, dtype=np.uint32)
        t = np.zeros(m, dtype=np.float32)
        bary = np.zeros((2, m), dtype=np.float32)
        st = np.zeros((m, 2), dtype=np.uint32) if stats else None
        if width == 4:
            nodes4 = self.nodes4()
            lib().om_trace4_closest(_ptr(self.tris), self.n, _ptr(self.sorted_idx), _ptr(nodes4), self.root4,
                                    _ptr(rays), m, m, tmin, _ptr(prim), _ptr(t), _ptr(bary), _ptr(st))
        else:
            lib().om_trace_closest(_ptr(self.tris), self.n, _ptr(self.sorted_idx), _ptr(self.left),
                                   _ptr(self.right), _ptr(self.node_aabb), self.root, _ptr(rays), m,
                                   m, tmin, _ptr(prim), _ptr(t), _ptr(bary), _ptr(st))
        if stats:
            return prim, t, bary, st
        return prim, t, bary

    def trace_any(self, rays: np.ndarray, tmin: float = 0.0, width: int = 4):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        m = rays.shape[1]
        out = np.zeros(m, dtype=np.uint32)
        if width == 4:
            nodes4 = self.nodes4()
            lib().om_trace4_any(_ptr(self.tris), self.n, _ptr(self.sorted_idx), _ptr(nodes4), self.root4,
                                _ptr(rays), m, m, tmin, _ptr(out))
        else:
            lib().om_trace_any(_ptr(self.tris), self.n, _ptr(self.sorted_idx), _ptr(self.left),
                               _ptr(self.right), _ptr(self.node_aabb), self.root, _ptr(rays), m, m,
                               tmin, _ptr(out))
        return out


def affine_inverse(l2w) -> np.ndarray:
    """World-to-local 3x4 of a row-major local-to-world 3x4 (double adjugate,
    rounded to float); raises on a singular matrix."""
    m = np.ascontiguousarray(l2w, np.float32).reshape(12)
    out = np.zeros(12, np.float32)
    if not lib().om_affine_inverse(_ptr(m), _ptr(out)):
        raise ValueError("singular transform")
    return out


def xf_ray(m12, o, d):
    m = np.ascontiguousarray(m12, np.float32).reshape(12)
    oo, dd = np.ascontiguousarray(o, np.float32), np.ascontiguousarray(d, np.float32)
    ro, rd = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().om_xf_ray(_ptr(m), _ptr(oo), _ptr(dd), _ptr(ro), _ptr(rd))
    return ro, rd


def world_box(l2w, box) -> np.ndarray:
    m = np.ascontiguousarray(l2w, np.float32).reshape(12)
    b = np.ascontiguousarray(box, np.float32).reshape(6)
    out = np.zeros(6, np.float32)
    lib().om_world_box(_ptr(m), _ptr(b), _ptr(out))
    return out


class TwoLevel:
    """Oracle two-level scene: one LBVH per mesh (local space), a top LBVH
    over the instances' world boxes given as degenerate triangles
    (min, max, max), world-to-local matrices by om_affine_inverse, and the
    light list (instance-major, each mesh in primitive order).
    scene["instancing"] = {"mesh_first": (m+1,), "xform": (k, 12),
    "inst_mesh": (k,), "inst_key": (k,) or absent}."""

    def __init__(self, scene: dict, bits: int = 30):
        inst = scene["instancing"]
        self.tris = np.ascontiguousarray(scene["tris"], np.float32).reshape(-1, 9)
        self.n = self.tris.shape[0]
        keys = np.ascontiguousarray(scene["prim_key"], np.uint32)
        first = np.ascontiguousarray(inst["mesh_first"], np.uint32)
        self.mesh_first = first
        nm = len(first) - 1
        self.l2w = np.ascontiguousarray(inst["xform"], np.float32).reshape(-1, 12)
        ni = self.l2w.shape[0]
        self.inst_mesh = np.ascontiguousarray(inst["inst_mesh"], np.uint32)
        ik = inst.get("inst_key")
        self.inst_key = None if ik is None else np.ascontiguousarray(ik, np.uint32)
        self.w2l = np.stack([affine_inverse(self.l2w[i]) for i in range(ni)]).astype(np.float32)
        self.meshes = [LBVH(self.tris[first[m]:first[m + 1]], bits) for m in range(nm)]
        self.mesh_sorted_idx = np.concatenate([b.sorted_idx for b in self.meshes]).astype(np.uint32)
        ib, ab = [], []
        i_acc = a_acc = 0
        for b in self.meshes:
            ib.append(i_acc)
            ab.append(a_acc)
            i_acc += b.n - 1
            a_acc += 2 * b.n - 1
        self.mesh_ibase = np.array(ib, np.uint32)
        self.mesh_abase = np.array(ab, np.uint32)
        cat = lambda xs, shape: (np.concatenate(xs) if sum(len(x) for x in xs) else np.zeros(shape))
        self.mesh_nodes4 = np.ascontiguousarray(cat([b.nodes4()[:b.n - 1] for b in self.meshes], (1, 32)),
                                                np.float32)
        self.mesh_left = np.ascontiguousarray(cat([b.left[:b.n - 1] for b in self.meshes], (1,)), np.uint32)
        self.mesh_right = np.ascontiguousarray(cat([b.right[:b.n - 1] for b in self.meshes], (1,)), np.uint32)
        self.mesh_node_aabb = np.ascontiguousarray(np.concatenate([b.node_aabb for b in self.meshes]), np.float32)
        self.mesh_root = np.array([b.root for b in self.meshes], np.uint32)
        self.mesh_root4 = np.array([b.root4 for b in self.meshes], np.uint32)
        # instance boxes: the mesh's vertex extent taken to the world
        boxes = np.zeros((ni, 9), np.float32)
        for i in range(ni):
            mt = self.tris[first[self.inst_mesh[i]]:first[self.inst_mesh[i] + 1]].reshape(-1, 3)
            box = np.concatenate([mt.min(0), mt.max(0)]).astype(np.float32)
            wb = world_box(self.l2w[i], box)
            boxes[i] = [wb[0], wb[1], wb[2], wb[3], wb[4], wb[5], wb[3], wb[4], wb[5]]
        self.boxes = boxes
        self.top = LBVH(boxes, bits)
        self.top.nodes4()
        lights, linst = [], []
        for i in range(ni):
            m = self.inst_mesh[i]
            f, e = int(first[m]), int(first[m + 1])
            ko = 0 if self.inst_key is None else int(self.inst_key[i])
            eff = np.full(e - f, ko, np.uint32) if ko else keys[f:e]
            for p in np.nonzero((eff >> 24) == 4)[0]:
                lights.append(f + int(p))
                linst.append(i)
        self.lights = np.array(lights, np.uint32)
        self.light_inst = np.array(linst, np.uint32)
        st = OmTwoLevel()
        st.n_inst = ni
        st.w2l, st.l2w = _ptr(self.w2l), _ptr(self.l2w)
        st.inst_mesh = _ptr(self.inst_mesh)
        st.inst_key = _ptr(self.inst_key)
        st.top_sorted_idx = _ptr(self.top.sorted_idx)
        st.top_nodes4 = _ptr(self.top.nodes4())
        st.top_root4 = self.top.root4
        st.n_meshes = nm
        st.mesh_first = _ptr(self.mesh_first)
        st.mesh_sorted_idx = _ptr(self.mesh_sorted_idx)
        st.mesh_nodes4 = _ptr(self.mesh_nodes4)
        st.mesh_left, st.mesh_right = _ptr(self.mesh_left), _ptr(self.mesh_right)
        st.mesh_node_aabb = _ptr(self.mesh_node_aabb)
        st.mesh_root, st.mesh_root4 = _ptr(self.mesh_root), _ptr(self.mesh_root4)
        st.mesh_ibase, st.mesh_abase = _ptr(self.mesh_ibase), _ptr(self.mesh_abase)
        self.struct = st

    def trace(self, rays: np.ndarray, tmin: float = 0.0, stats: bool = False, width: int = 4):
        """Closest hits: (inst, prim, t, bary[2, m]) (+ stats [m, 2])."""
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        m = rays.shape[1]
        inst = np.zeros(m, np.uint32)
        prim = np.zeros(m, np.uint32)
        t = np.zeros(m, np.float32)
        bary = np.zeros((2, m), np.float32)
        st = np.zeros((m, 2), np.uint32) if stats else None
        lib().om_trace2l_closest(C.byref(self.struct), _ptr(self.tris), width, _ptr(rays), m, m, tmin, _ptr(inst),
                                 _ptr(prim), _ptr(t), _ptr(bary), _ptr(st))
        if stats:
            return inst, prim, t, bary, st
        return inst, prim, t, bary


def trace_brute(tris: np.ndarray, rays: np.ndarray, tmin: float = 0.0):
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    rays = np.ascontiguousarray(rays, dtype=np.float32)
    m = rays.shape[1]
    prim = np.zeros(m, dtype=np.uint32)
    t = np.zeros(m, dtype=np.float32)
    bary = np.zeros((2, m), dtype=np.float32)
    lib().om_trace_brute(_ptr(tris), tris.shape[0], _ptr(rays), m, m, tmin, _ptr(prim),
                         _ptr(t), _ptr(bary))
    return prim, t, bary


def camera(pos, gaze, up, fov_x, near, far, w, h) -> OmCamera:
    cam = OmCamera()
    f3 = lambda v: np.ascontiguousarray(v, dtype=np.float32)
    p, g, u = f3(pos), f3(gaze), f3(up)
    lib().om_camera_setup(_ptr(p), _ptr(g), _ptr(u), fov_x, near, far, w, h, C.byref(cam))
    return cam


def camera_ray(cam: OmCamera, w, h, samples, seed, sample_id):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    st = C.c_uint64(0)
    lib().om_camera_ray(C.byref(cam), w, h, samples, seed, sample_id, _ptr(o), _ptr(d),
                        C.byref(st))
    return o, d, st.value


def camera_ray_at(cam: OmCamera, w, h, samples, sample_id, jx, jy):
    """Camera ray of one sample with a given stratum jitter (jx, jy)."""
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().om_camera_ray_at(C.byref(cam), w, h, samples, sample_id, jx, jy, _ptr(o), _ptr(d))
    return o, d


def pcg32_stream(initstate: int, initseq: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().om_pcg32_stream(initstate, initseq, n, _ptr(out))
    return out


def splitmix64(z: int) -> int:
    return int(lib().om_splitmix64(z))


def path_uniforms(seed: int, sample_id: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float32)
    lib().om_path_uniforms(seed, sample_id, n, _ptr(out))
    return out


def corner_frames(normals) -> np.ndarray:
    """Per-corner world-to-tangent rotations [n][3][4] (w, x, y, z) of
    per-corner shading normals [n][9] (om_corner_frames)."""
    nrm = np.ascontiguousarray(normals, np.float32).reshape(-1, 9)
    out = np.zeros((nrm.shape[0], 3, 4), np.float32)
    lib().om_corner_frames(_ptr(nrm), nrm.shape[0], _ptr(out))
    return out


def _scene_struct(scene: dict, bvh: LBVH, keep: dict, width: int = 4) -> OmScene:

    def arr(name, dtype, cols):
        a = scene.get(name)
        if a is None or len(a) == 0:
            a = np.zeros((1, cols), dtype=dtype)
        a = np.ascontiguousarray(a, dtype=dtype)
        keep[name] = a
        return _ptr(a)

    sc = OmScene()
    sc.tris = arr("tris", np.float32, 9)
    sc.prim_key = arr("prim_key", np.uint32, 1)
    sc.n = bvh.n
    sc.lambert = arr("lambert", np.float32, 3)
    sc.metal = arr("metal", np.float32, 10)
    sc.dielectric = arr("dielectric", np.float32, 4)
    sc.emissive = arr("emissive", np.float32, 3)
    sc.reflect = arr("reflect", np.float32, 4)
    if isinstance(bvh, TwoLevel):
        keep["two_level"] = bvh.struct
        sc.two_level = C.cast(C.pointer(bvh.struct), _P)
        sc.tree_width = width
        sc.n = bvh.n
        keep["lights"] = bvh.lights if len(bvh.lights) else np.zeros(1, np.uint32)
        keep["light_inst"] = bvh.light_inst if len(bvh.light_inst) else np.zeros(1, np.uint32)
        sc.lights = _ptr(keep["lights"])
        sc.light_inst = _ptr(keep["light_inst"])
        sc.n_lights = len(bvh.lights)
        return sc
    if scene.get("normals") is not None:
        keep["tbn"] = corner_frames(scene["normals"])
        sc.tbn = _ptr(keep["tbn"])
    sc.sorted_idx = _ptr(bvh.sorted_idx)
    sc.left = _ptr(bvh.left)
    sc.right = _ptr(bvh.right)
    sc.node_aabb = _ptr(bvh.node_aabb)
    sc.root = bvh.root
    if width == 4:
        sc.nodes4 = _ptr(bvh.nodes4())
        sc.root4 = bvh.root4
    keys = np.ascontiguousarray(scene["prim_key"], dtype=np.uint32)
    lights = np.nonzero((keys >> 24) == 4)[0].astype(np.uint32)
    keep["lights"] = lights if len(lights) else np.zeros(1, np.uint32)
    sc.lights = _ptr(keep["lights"])
    sc.n_lights = len(lights)
    return sc


def render(scene: dict, bvh: LBVH, cam: OmCamera, width, height, samples, max_depth, rr_start,
           seed, row_step=1, row_offset=0, threads=1, tree_width=4, nee=0, mis=0):
    """scene: dict with tris (n,9), prim_key (n,), lambert (.,3), metal (.,10),
    dielectric (.,4), emissive (.,3); tree_width = traversal tree (2 or 4);
    nee / mis = next-event estimation / its MIS with the BxDF.
    Returns (image float64 [h,w,3], segments)."""
    keep = {}
    sc = _scene_struct(scene, bvh, keep, tree_width)
    p = OmRenderParams(width, height, samples, max_depth, rr_start, seed, row_step, row_offset,
                       threads, nee, mis)
    img = np.zeros((height, width, 3), dtype=np.float64)
    seg = lib().om_render(C.byref(sc), C.byref(cam), C.byref(p), _ptr(img))
    return img / float(samples * samples), int(seg)


def filter_pixel_wh(radius: float) -> int:
    return int(lib().om_filter_pixel_wh(float(radius)))


def filter_weight(desc: OmFilterDesc, dx: float, dy: float) -> float:
    return float(lib().om_filter_weight(C.byref(desc), float(dx), float(dy)))


def filter_samples(desc: OmFilterDesc, values, coords, w: int, h: int, grouped_spp: int = 0):
    """Filter samples (values n x 3, coords n x 2) into (weighted h x w x 3,
    weights h x w); grouped_spp > 0: sample i lies in pixel i // grouped_spp."""
    v = np.ascontiguousarray(values, np.float32)
    c = np.ascontiguousarray(coords, np.float32)
    weighted = np.zeros((h, w, 3), np.float32)
    weights = np.zeros((h, w), np.float32)
    if grouped_spp:
        lib().om_filter_samples_grouped(C.byref(desc), _ptr(v), _ptr(c), grouped_spp, w, h, _ptr(weighted),
                                        _ptr(weights))
    else:
        lib().om_filter_samples(C.byref(desc), _ptr(v), _ptr(c), len(v), w, h, _ptr(weighted), _ptr(weights))
    return weighted, weights


def filter_resolve(weighted, weights):
    ws = weights[..., None]
    with np.errstate(divide="ignore", invalid="ignore"):
        out = np.where(ws == 0, np.float32(0), weighted / np.where(ws == 0, np.float32(1), ws))
    return out.astype(np.float32)


def render_samples(scene: dict, bvh: "LBVH", cam: OmCamera, width, height, samples, max_depth, rr_start, seed,
                   threads=1, tree_width=4, nee=0, mis=0, sample_base=0):
    """Every sample's FP32 radiance and image coordinate:
    (values [n, 3], coords [n, 2], segments), n = width * height * samples^2;
    global sample ids start at sample_base."""
    keep = {}
    sc = _scene_struct(scene, bvh, keep, tree_width)
    p = OmRenderParams(width, height, samples, max_depth, rr_start, seed, 1, 0, threads, nee, mis)
    n = width * height * samples * samples
    values = np.zeros((n, 3), np.float32)
    coords = np.zeros((n, 2), np.float32)
    seg = lib().om_render_samples(C.byref(sc), C.byref(cam), C.byref(p), sample_base, _ptr(values), _ptr(coords))
    return values, coords, int(seg)


def render_filtered(scene: dict, bvh: "LBVH", cam: OmCamera, width, height, samples, max_depth, rr_start, seed,
                    desc: OmFilterDesc, threads=1, tree_width=4, nee=0, mis=0):
    """Filtered frame: (image float32 [h, w, 3], segments)."""
    values, coords, seg = render_samples(scene, bvh, cam, width, height, samples, max_depth, rr_start, seed,
                                         threads, tree_width, nee, mis)
    weighted, weights = filter_samples(desc, values, coords, width, height, samples * samples)
    return filter_resolve(weighted, weights), seg


def sample_light(scene: dict, bvh: "LBVH", p, xi):
    """Light sample of the next-event estimator from point p with numbers xi
    (light choice, two barycentric numbers): (light primitive, unit direction,
    squared distance, pdf = solid-angle pdf / light count)."""
    keep = {}
    sc = _scene_struct(scene, bvh, keep, 2)
    pp = np.ascontiguousarray(p, np.float32)
    xx = np.ascontiguousarray(xi, np.float32)
    d = np.zeros(3, np.float32)
    d2, pdf = _F(0), _F(0)
    prim = lib().om_sample_light(C.byref(sc), _ptr(pp), _ptr(xx), _ptr(d), C.byref(d2), C.byref(pdf))
    return int(prim), d, float(d2.value), float(pdf.value)


def light_pdf_hit(tri, t, d) -> float:
    """Solid-angle pdf of reaching triangle tri (9 floats) at distance t along unit d."""
    tt = np.ascontiguousarray(tri, np.float32).reshape(9)
    dd = np.ascontiguousarray(d, np.float32)
    return float(lib().om_light_pdf_hit(_ptr(tt), float(t), _ptr(dd)))


def power_heuristic(pdf0, pdf1) -> float:
    return float(lib().om_power_heuristic(float(pdf0), float(pdf1)))


class Svo:
    """The oracle's guide octree (om_svo), with its arrays copied out."""

    def __init__(self, s: OmSvo):
        self._s = s
        self.depth = s.depth
        self.nodes = s.nodes
        self.cube = np.array(s.cube[:], dtype=np.float32)
        self.level_count = np.array(s.level_count[:], dtype=np.uint32)
        self.level_offset = np.array(s.level_offset[:], dtype=np.uint32)
        self.refresh()

    def refresh(self):
        n = self.nodes

        def get(ptr, ctype, count, dtype):
            return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(count,)).astype(dtype).copy()

        self.codes = get(self._s.codes, C.c_uint32, n, np.uint32)
        self.first_child = get(self._s.first_child, C.c_uint32, n, np.uint32)
        self.child_mask = get(self._s.child_mask, C.c_uint32, n, np.uint32)
        self.parent = get(self._s.parent, C.c_uint32, n, np.uint32)
        self.acc = get(self._s.acc, C.c_uint64, n * 64, np.uint64).reshape(n, 64)
        self.count = get(self._s.count, C.c_uint32, n * 64, np.uint32).reshape(n, 64)
        self.sum = get(self._s.sum, C.c_float, n * 64, np.float32).reshape(n, 64)
        self.radiance = get(self._s.radiance, C.c_float, n * 64, np.float32).reshape(n, 64)
        self.node_count = get(self._s.node_count, C.c_uint32, n, np.uint32)
        self.centers = np.array(self._s.centers[:], dtype=np.float32).reshape(64, 3)

    @classmethod
    def build(cls, codes: np.ndarray, cube, depth: int) -> "Svo":
        codes = np.ascontiguousarray(codes, dtype=np.uint32)
        cube = np.ascontiguousarray(cube, dtype=np.float32)
        s = OmSvo()
        rc = lib().om_svo_build(C.byref(s), _ptr(codes), len(codes), _ptr(cube), depth)
        if rc != 0:
            raise ValueError(f"om_svo_build failed ({rc})")
        return cls(s)

    def cone_trace(self, rays: np.ndarray, aperture: float, max_query_level_offset: int = 0):
        """om_svo_cone_trace over rays [n][7] (origin, direction, tmin):
        (nodes, t, is_leaf)."""
        rays = np.ascontiguousarray(rays, dtype=np.float32).reshape(-1, 7)
        n = len(rays)
        nodes = np.zeros(n, np.uint32)
        t = np.zeros(n, np.float32)
        leaf = np.zeros(n, np.uint32)
        lib().om_svo_cone_trace_batch(C.byref(self._s), _ptr(rays), n, aperture, max_query_level_offset,
                                      _ptr(nodes), _ptr(t), _ptr(leaf))
        return nodes, t, leaf

    def payload(self, tris, codes, prims, vertices):
        """om_svo_payload: (irr [nodes][2], normals [nodes][4])."""
        tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
        codes = np.ascontiguousarray(codes, dtype=np.uint32)
        prims = np.ascontiguousarray(prims, dtype=np.uint32)
        vertices = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 7)
        irr = np.zeros((self.nodes, 2), np.float32)
        nrm = np.zeros((self.nodes, 4), np.float32)
        lib().om_svo_payload(C.byref(self._s), _ptr(tris), _ptr(codes), _ptr(prims), len(codes), _ptr(vertices),
                             len(vertices), _ptr(irr), _ptr(nrm))
        return irr, nrm

    def trace_fields(self, X: int, Y: int, origins, jitter, irr, normals, boundary: float = 0.0, offset: int = 0):
        """om_wfpg_trace_fields: (fields [bins][Y][X], nodes, t)."""
        origins = np.ascontiguousarray(origins, dtype=np.float32).reshape(-1, 4)
        jitter = np.ascontiguousarray(jitter, dtype=np.float32).reshape(-1, 2)
        irr = np.ascontiguousarray(irr, dtype=np.float32).reshape(-1, 2)
        normals = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
        bins = len(origins)
        fields = np.zeros((bins, Y, X), np.float32)
        nodes = np.zeros((bins, Y, X), np.uint32)
        t = np.zeros((bins, Y, X), np.float32)
        lib().om_wfpg_trace_fields(C.byref(self._s), X, Y, bins, _ptr(origins), _ptr(jitter), _ptr(irr),
                                   _ptr(normals), boundary, offset, _ptr(fields), _ptr(nodes), _ptr(t))
        return fields, nodes, t

    def lookup(self, codes: np.ndarray) -> np.ndarray:
        return np.array([lib().om_svo_lookup(C.byref(self._s), int(c)) for c in codes], dtype=np.uint32)

    def lookup_guide(self, codes: np.ndarray, min_count: int) -> np.ndarray:
        return np.array([lib().om_svo_lookup_guide(C.byref(self._s), int(c), min_count) for c in codes],
                        dtype=np.uint32)

    def splat(self, codes, bins, values):
        codes = np.ascontiguousarray(codes, dtype=np.uint32)
        bins = np.ascontiguousarray(bins, dtype=np.uint32)
        values = np.ascontiguousarray(values, dtype=np.float32)
        lib().om_svo_splat(C.byref(self._s), _ptr(codes), _ptr(bins), _ptr(values), len(codes))
        self.refresh()

    def filter(self):
        lib().om_svo_filter(C.byref(self._s))
        self.refresh()

    def bin_rays(self, codes, min_level: int, min_rays: int):
        """om_svo_bin_rays: (bins, order, part_offsets[count + 1], part_bins[count])."""
        codes = np.ascontiguousarray(codes, dtype=np.uint32)
        n = len(codes)
        bins = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        offsets = np.zeros(n + 1, np.uint32)
        pbins = np.zeros(max(n, 1), np.uint32)
        count = np.zeros(1, np.uint32)
        lib().om_svo_bin_rays(C.byref(self._s), _ptr(codes), n, min_level, min_rays, _ptr(bins), _ptr(order),
                              _ptr(offsets), _ptr(pbins), _ptr(count))
        k = int(count[0])
        return bins, order, offsets[:k + 1].copy(), pbins[:k].copy()

    def sample(self, nodes, normals, xi):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        normals = np.ascontiguousarray(normals, dtype=np.float32)
        xi = np.ascontiguousarray(xi, dtype=np.float32)
        n = len(nodes)
        dirs = np.zeros((n, 3), dtype=np.float32)
        pdf = np.zeros(n, dtype=np.float32)
        lib().om_guide_sample(C.byref(self._s), _ptr(nodes), _ptr(normals), _ptr(xi), n, _ptr(dirs), _ptr(pdf))
        return dirs, pdf

    def pdf(self, nodes, normals, dirs):
        nodes = np.ascontiguousarray(nodes, dtype=np.uint32)
        normals = np.ascontiguousarray(normals, dtype=np.float32)
        dirs = np.ascontiguousarray(dirs, dtype=np.float32)
        n = len(nodes)
        pdf = np.zeros(n, dtype=np.float32)
        lib().om_guide_pdf(C.byref(self._s), _ptr(nodes), _ptr(normals), _ptr(dirs), n, _ptr(pdf))
        return pdf

    def close(self):
        if self._s is not None:
            lib().om_svo_free(C.byref(self._s))
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def svo_codes(pts: np.ndarray, cube, depth: int) -> np.ndarray:
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    cube = np.ascontiguousarray(cube, dtype=np.float32)
    out = np.zeros(len(pts), dtype=np.uint32)
    lib().om_svo_codes(_ptr(pts), len(pts), _ptr(cube), depth, _ptr(out))
    return out


def dir_to_bins(dirs: np.ndarray) -> np.ndarray:
    dirs = np.ascontiguousarray(dirs, dtype=np.float32)
    out = np.zeros(len(dirs), dtype=np.uint32)
    lib().om_dir_to_bins(_ptr(dirs), len(dirs), _ptr(out))
    return out


def octa_to_dirs(uv: np.ndarray) -> np.ndarray:
    uv = np.ascontiguousarray(uv, dtype=np.float32)
    out = np.zeros((len(uv), 3), dtype=np.float32)
    lib().om_octa_to_dirs(_ptr(uv), len(uv), _ptr(out))
    return out


def dir_to_octas(dirs: np.ndarray) -> np.ndarray:
    dirs = np.ascontiguousarray(dirs, dtype=np.float32)
    out = np.zeros((len(dirs), 2), dtype=np.float32)
    lib().om_dir_to_octas(_ptr(dirs), len(dirs), _ptr(out))
    return out


def atan_unit(t: float) -> float:
    return float(lib().om_atan_unit(t))


def render_guided(scene: dict, bvh: LBVH, cam: OmCamera, width, height, samples, max_depth, rr_start,
                  seed, iterations, svo_depth, guide_ratio=0.5, min_samples=0, threads=1, keep_svo=False,
                  tree_width=4, nee=0, mis=0):
    """Guided render of every row.  Returns (image float64 [h,w,3] = mean over
    all frames, segments, Svo of the last guide or None)."""
    keep = {}
    sc = _scene_struct(scene, bvh, keep, tree_width)
    p = OmRenderParams(width, height, samples, max_depth, rr_start, seed, 1, 0, threads, nee, mis)
    g = OmGuideParams(iterations, svo_depth, guide_ratio, min_samples)
    img = np.zeros((height, width, 3), dtype=np.float64)
    last = OmSvo()
    seg = lib().om_render_guided(C.byref(sc), C.byref(cam), C.byref(p), C.byref(g), _ptr(img),
                                 C.byref(last))
    svo = None
    if last.nodes:
        if keep_svo:
            svo = Svo(last)
        else:
            lib().om_svo_free(C.byref(last))
    return img / float(samples * samples * iterations), int(seg), svo


# ---- per-bin radiance fields of the reference's wavefront path guiding

def wfpg_gauss_weights(alpha: float) -> np.ndarray:
    w = np.zeros(11, np.float32)
    lib().om_wfpg_gauss_weights(alpha, _ptr(w))
    return w


def octa_wrap_int(x: int, y: int, X: int, Y: int) -> tuple[int, int]:
    out = np.zeros(2, np.int32)
    lib().om_octa_wrap_int(x, y, X, Y, _ptr(out))
    return int(out[0]), int(out[1])


def wfpg_filter(field: np.ndarray, alpha: float) -> np.ndarray:
    """field [Y][X] -> Gaussian-filtered field (no empty-texel raise)."""
    field = np.ascontiguousarray(field, dtype=np.float32)
    Y, X = field.shape
    out = np.zeros_like(field)
    lib().om_wfpg_filter(_ptr(field), X, Y, _ptr(wfpg_gauss_weights(alpha)), _ptr(out))
    return out


class Pwc2D:
    """om_wfpg_pwc over a [Y][X] field: pdfx/cdfx [Y][X], pdfy/cdfy [Y]."""

    def __init__(self, f: np.ndarray):
        f = np.ascontiguousarray(f, dtype=np.float32)
        self.Y, self.X = f.shape
        self.pdfx = np.zeros_like(f)
        self.cdfx = np.zeros_like(f)
        self.pdfy = np.zeros(self.Y, np.float32)
        self.cdfy = np.zeros(self.Y, np.float32)
        lib().om_wfpg_pwc(_ptr(f), self.X, self.Y, _ptr(self.pdfx), _ptr(self.cdfx), _ptr(self.pdfy),
                          _ptr(self.cdfy))

    def sample(self, xi0: float, xi1: float) -> np.ndarray:
        uv = np.zeros(2, np.float32)
        lib().om_wfpg_pwc_sample(_ptr(self.cdfx), _ptr(self.cdfy), self.X, self.Y, xi0, xi1, _ptr(uv))
        return uv

    def pdf(self, u: float, v: float) -> float:
        return float(lib().om_wfpg_pwc_pdf(_ptr(self.pdfx), _ptr(self.pdfy), self.X, self.Y, u, v))


def wfpg_sample_fields(fields: np.ndarray, bin_offsets: np.ndarray, ray_ids: np.ndarray, xi: np.ndarray,
                       alpha: float = 1.0, eval_uv: np.ndarray | None = None, ray_count: int | None = None,
                       bxdf: np.ndarray | None = None, seed: int = 0):
    """om_wfpg_sample_fields.  fields [bins][Y][X].  Returns dict of uv [n][2],
    dir [n][3], pdf [n], eval_pdf [n] (or None), filtered [bins][Y][X]; rays
    outside every bin keep NaN."""
    fields = np.ascontiguousarray(fields, dtype=np.float32)
    bins, Y, X = fields.shape
    bin_offsets = np.ascontiguousarray(bin_offsets, dtype=np.uint32)
    ray_ids = np.ascontiguousarray(ray_ids, dtype=np.uint32)
    xi = np.ascontiguousarray(xi, dtype=np.float32)
    n = len(xi) if ray_count is None else ray_count
    out = {"uv": np.full((n, 2), np.nan, np.float32), "dir": np.full((n, 3), np.nan, np.float32),
           "pdf": np.full(n, np.nan, np.float32), "filtered": np.zeros_like(fields), "eval_pdf": None}
    ev = None
    if eval_uv is not None:
        ev = np.ascontiguousarray(eval_uv, dtype=np.float32)
        out["eval_pdf"] = np.full(n, np.nan, np.float32)
    bx = np.ascontiguousarray(bxdf, dtype=np.float32) if bxdf is not None else None
    d = OmWfpgDesc(X, Y, bins, n, alpha, _ptr(fields), _ptr(bin_offsets), _ptr(ray_ids), _ptr(xi), _ptr(ev),
                   _ptr(out["uv"]), _ptr(out["dir"]), _ptr(out["pdf"]), _ptr(out["eval_pdf"]),
                   _ptr(out["filtered"]), _ptr(bx), seed)
    lib().om_wfpg_sample_fields(C.byref(d))
    return out


def voxelize(tris: np.ndarray, cube, depth: int, prims: bool = False):
    """om_voxelize: voxel cell codes of every covered pixel (duplicates kept);
    with prims=True also each record's triangle index."""
    tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
    cube = np.ascontiguousarray(cube, dtype=np.float32)
    n = int(lib().om_voxelize(_ptr(tris), len(tris), _ptr(cube), depth, None, None, 0))
    out = np.zeros(max(n, 1), np.uint32)
    pr = np.zeros(max(n, 1), np.uint32)
    lib().om_voxelize(_ptr(tris), len(tris), _ptr(cube), depth, _ptr(out), _ptr(pr), n)
    return (out[:n], pr[:n]) if prims else out[:n]


def svo_cube(lo, hi, depth: int) -> np.ndarray:
    """The guiding octree's cube for scene bounds [lo, hi] at `depth`
    (AnisoSVOctreeCPU::Constrcut, AnisoSVO.cu:670-678): the largest span,
    padded by one leaf on every side."""
    lo = np.asarray(lo, np.float32)
    span = np.asarray(hi, np.float32) - lo
    w = np.float32(span.max())
    pad = np.float32(w / np.float32(1 << depth))
    return np.array([lo[0] - pad, lo[1] - pad, lo[2] - pad, np.float32(w + np.float32(2) * pad)], np.float32)

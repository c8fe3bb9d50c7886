"""ctypes mirror of include/tempest_rt.h and include/tempest_rt_debug.h.

The structures keep the C field order exactly; `load()` opens the HIP
library built in-tree (libtempest_rt.so next to this file) and declares
every exported entry point.  A missing library raises: the tracer has no
CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtempest_rt.so")

c_float3 = ctypes.c_float * 3


class tr_material(ctypes.Structure):
    _fields_ = [
        ("model", ctypes.c_uint32),
        ("specular", ctypes.c_float * 3),
        ("specular_power", ctypes.c_float * 2),
        ("diffuse", ctypes.c_float * 3),
        ("fresnel", ctypes.c_float * 2),
        ("normalization", ctypes.c_float),
        ("sggx_basis", ctypes.c_float * 4),
        ("sggx_depth", ctypes.c_uint32),
        ("sggx_sample_count", ctypes.c_uint32),
        ("albedo", ctypes.c_float * 3),
        ("sggx_stddev", ctypes.c_float * 3),
        ("diffuse_map", ctypes.c_uint32),
        ("specular_map", ctypes.c_uint32),
        ("stddev_map", ctypes.c_uint32),
        ("basis_map", ctypes.c_uint32),
    ]


class tr_light(ctypes.Structure):
    _fields_ = [
        ("type", ctypes.c_uint32),
        ("direction", ctypes.c_float * 3),
        ("position", ctypes.c_float * 3),
        ("radiance", ctypes.c_float * 3),
        ("sphere_center", ctypes.c_float * 3),
        ("sphere_radius", ctypes.c_float),
        ("geometry_id", ctypes.c_uint32),
    ]


class tr_submesh(ctypes.Structure):
    _fields_ = [
        ("material", ctypes.POINTER(tr_material)),
        ("vertex_count", ctypes.c_uint32),
        ("vertex_offset", ctypes.c_uint32),
        ("base_index", ctypes.c_uint32),
        ("stride", ctypes.c_uint32),
    ]


class tr_sample_data(ctypes.Structure):
    _fields_ = [
        ("material", ctypes.c_void_p),
        ("normal", ctypes.c_float * 3),
        ("tangent", ctypes.c_float * 3),
        ("binormal", ctypes.c_float * 3),
        ("texcoord", ctypes.c_float * 2),
        ("t", ctypes.c_float),
    ]


class tr_mesh_options(ctypes.Structure):
    _fields_ = [("two_sided", ctypes.c_int32)]


class tr_settings(ctypes.Structure):
    _fields_ = [("data_pool_size", ctypes.c_uint32), ("path_pool_size", ctypes.c_uint32),
                ("trace_waves_per_cu", ctypes.c_uint32), ("shade_order", ctypes.c_uint32)]


class tr_ray(ctypes.Structure):
    _fields_ = [("org", ctypes.c_float * 3), ("tnear", ctypes.c_float),
                ("dir", ctypes.c_float * 3), ("tfar", ctypes.c_float)]


class tr_hit(ctypes.Structure):
    _fields_ = [("t", ctypes.c_float), ("u", ctypes.c_float), ("v", ctypes.c_float),
                ("geom", ctypes.c_uint32), ("prim", ctypes.c_uint32)]


class tr_stats(ctypes.Structure):
    _fields_ = [
        ("paths", ctypes.c_uint64), ("segments", ctypes.c_uint64), ("shadow_rays", ctypes.c_uint64),
        ("frame_ms", ctypes.c_double), ("trace_ms", ctypes.c_double), ("shade_ms", ctypes.c_double),
        ("shadow_ms", ctypes.c_double), ("other_ms", ctypes.c_double), ("trace_launches", ctypes.c_uint64),
        ("node_visits", ctypes.c_uint64), ("prim_tests", ctypes.c_uint64), ("counted_rays", ctypes.c_uint64),
        ("volume_steps", ctypes.c_uint64), ("cache_ms", ctypes.c_double), ("cache_loss", ctypes.c_double),
        ("exposure", ctypes.c_double),
    ]


class tr_image(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_uint32),
                ("data", ctypes.POINTER(ctypes.c_uint8))]


class tr_mesh_blob(ctypes.Structure):
    _fields_ = [("submesh_count", ctypes.c_uint32), ("submeshes", ctypes.POINTER(tr_submesh)),
                ("vertex_data", ctypes.POINTER(ctypes.c_uint8)), ("vertex_bytes", ctypes.c_size_t),
                ("index_data", ctypes.POINTER(ctypes.c_int32)), ("index_count", ctypes.c_size_t),
                ("stride", ctypes.c_uint32), ("materials", ctypes.POINTER(tr_material)),
                ("material_count", ctypes.c_uint32), ("ignored_maps", ctypes.c_uint32),
                ("textures", ctypes.POINTER(tr_image)), ("texture_count", ctypes.c_uint32)]


TR_OBJ_GENERATE_TANGENTS = 1
TR_OBJ_GENERATE_CONSISTENT_NORMALS = 2


class tr_volume_brick(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32 * 3), ("data", ctypes.POINTER(ctypes.c_uint8))]


class tr_volume_root(ctypes.Structure):
    _fields_ = [("min_corner", ctypes.c_float * 3), ("max_corner", ctypes.c_float * 3), ("dims", ctypes.c_int32 * 3),
                ("bricks", ctypes.POINTER(tr_volume_brick))]


class tr_texture_desc(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("format", ctypes.c_uint32),
                ("sampling", ctypes.c_uint32)]


TR_TEX_R8_UNORM, TR_TEX_RG8_UNORM, TR_TEX_RGBA8_UNORM, TR_TEX_R32F, TR_TEX_RG32F, TR_TEX_RGB32F, TR_TEX_RGBA32F = range(7)
TR_TEX_BILINEAR, TR_TEX_NEAREST = 0, 1


class tr_request(ctypes.Structure):
    _fields_ = [("view_proj_inv", ctypes.c_float * 16), ("width", ctypes.c_uint32), ("height", ctypes.c_uint32),
                ("samples_camera", ctypes.c_uint32), ("light_direction", ctypes.c_float * 3)]


class tr_request_options(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_int32), ("autotune", ctypes.c_int32), ("pool_slots", ctypes.c_uint32),
                ("trace_waves_per_cu", ctypes.c_uint32)]


class tr_request_stats(ctypes.Structure):
    _fields_ = [("paths", ctypes.c_uint64), ("pixels", ctypes.c_uint64), ("batches", ctypes.c_uint32),
                ("pool_slots", ctypes.c_uint32), ("trace_waves_per_cu", ctypes.c_uint32),
                ("candidates", ctypes.c_uint32), ("ms", ctypes.c_double), ("tune_ms", ctypes.c_double)]


class tr_cache_settings(ctypes.Structure):
    _fields_ = [("enabled", ctypes.c_int32), ("train", ctypes.c_int32), ("train_paths", ctypes.c_uint32),
                ("batch", ctypes.c_uint32), ("rough_alpha", ctypes.c_float), ("query_nee", ctypes.c_int32),
                ("seed", ctypes.c_uint64)]


class tr_grid_volume(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_uint32 * 3),
        ("min_corner", ctypes.c_float * 3),
        ("max_corner", ctypes.c_float * 3),
        ("density", ctypes.POINTER(ctypes.c_uint16)),
        ("direction", ctypes.POINTER(ctypes.c_uint16)),
        ("max_density", ctypes.c_float),
    ]


class tr_scene_info(ctypes.Structure):
    _fields_ = [
        ("primitives", ctypes.c_uint64), ("bvh_nodes", ctypes.c_uint64), ("bvh_bytes", ctypes.c_uint64),
        ("prim_bytes", ctypes.c_uint64), ("bvh_depth", ctypes.c_uint32), ("geometries", ctypes.c_uint32),
        ("lights", ctypes.c_uint32), ("materials", ctypes.c_uint32), ("build_ms", ctypes.c_double),
    ]


# numeric constants of tempest_rt.h
TR_OK = 0
TR_ERROR_INVALID_ARG = 1
TR_ERROR_OUT_OF_MEMORY = 2
TR_ERROR_DEVICE = 3
TR_ERROR_NOT_COMMITTED = 4
TR_ERROR_UNSUPPORTED = 5
TR_ERROR_IO = 6
TR_IMAGE_TGA = 0
TR_IMAGE_PNG = 1
TR_INVALID_GEOMETRY = 0xFFFFFFFF

P = ctypes.POINTER
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_st = ctypes.c_int

# name -> (restype, argtypes)
SIGNATURES = {
    "tr_last_error_message": (ctypes.c_char_p, []),
    "tr_bench_pseudo_volume_image": (_st, [P(tr_material), _u32, _u32, _u32, P(_u32), P(ctypes.c_float),
                                           P(ctypes.c_float)]),
    "tr_material_init": (_st, [P(tr_material), _u32]),
    "tr_material_setup": (_st, [P(tr_material)]),
    "tr_scene_create": (_st, [_u32, _u32, P(ctypes.c_float), P(tr_settings), P(_vp)]),
    "tr_scene_destroy": (None, [_vp]),
    "tr_scene_set_chunk_size": (_st, [_vp, _u32]),
    "tr_scene_set_samples_local_area_light": (_st, [_vp, _u32]),
    "tr_scene_set_samples_global_illumination": (_st, [_vp, _u32]),
    "tr_scene_set_samples_camera": (_st, [_vp, _u32]),
    "tr_scene_set_max_ray_depth": (_st, [_vp, _u32]),
    "tr_scene_set_russian_roulette": (_st, [_vp, ctypes.c_float]),
    "tr_scene_set_background_spectrum": (_st, [_vp, P(ctypes.c_float)]),
    "tr_scene_set_picture_post_process": (_st, [_vp, _u32]),
    "tr_scene_set_transparent_background": (_st, [_vp, ctypes.c_int32]),
    "tr_scene_set_render_mode": (_st, [_vp, _u32]),
    "tr_scene_repaint": (_st, [_vp]),
    "tr_scene_add_light_source": (_st, [_vp, P(tr_light), P(_u32)]),
    "tr_scene_update_light_source": (_st, [_vp, _u32, P(tr_light)]),
    "tr_scene_clear_light_sources": (_st, [_vp]),
    "tr_scene_add_sphere_light_source": (_st, [_vp, P(tr_light), P(_u32)]),
    "tr_scene_add_triangle_mesh": (_st, [_vp, P(ctypes.c_float), _sz, P(tr_submesh), _sz, P(ctypes.c_int32),
                                         _sz, _vp, P(tr_mesh_options), P(_u64)]),
    "tr_scene_add_sphere": (_st, [_vp, P(ctypes.c_float), ctypes.c_float, P(tr_material), P(_u64)]),
    "tr_scene_add_rect": (_st, [_vp, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                P(tr_material), P(ctypes.c_float), _u32, P(_u64)]),
    "tr_scene_add_blocker": (_st, [_vp, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                   P(_u64)]),
    "tr_scene_add_disk": (_st, [_vp, P(ctypes.c_float), P(ctypes.c_float), ctypes.c_float, ctypes.c_float,
                                P(tr_material), P(_u64)]),
    "tr_scene_add_cylinder": (_st, [_vp, P(ctypes.c_float), ctypes.c_float, ctypes.c_float, P(tr_material), P(_u64)]),
    "tr_scene_add_oblique_cylinder": (_st, [_vp, P(ctypes.c_float), ctypes.c_float, ctypes.c_float, P(ctypes.c_float),
                                            ctypes.c_float, P(tr_material), P(_u64)]),
    "tr_scene_add_ellipsoid": (_st, [_vp, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float), P(tr_material),
                                     P(_u64)]),
    "tr_scene_add_grid_volume": (_st, [_vp, P(tr_grid_volume), P(tr_material), P(_u64)]),
    "tr_scene_commit": (_st, [_vp]),
    "tr_scene_get_info": (_st, [_vp, P(tr_scene_info)]),
    "tr_scene_draw_once": (_st, [_vp]),
    "tr_scene_draw": (_st, [_vp, _u32, _u32, P(ctypes.c_float)]),
    "tr_scene_frame_wait": (_st, [_vp]),
    "tr_scene_read_backbuffer": (_st, [_vp, P(_u32), _sz]),
    "tr_scene_read_accum": (_st, [_vp, P(ctypes.c_double), _sz, P(_u64)]),
    "tr_scene_bind_texture": (_st, [_vp, P(tr_texture_desc), _vp, P(_u32)]),
    "tr_scene_set_global_cube_map": (_st, [_vp, _u32, P(_vp)]),
    "tr_parse_volume_hierarchy": (_st, [ctypes.c_char_p, P(tr_volume_root)]),
    "tr_parse_volume": (_st, [ctypes.c_char_p, ctypes.c_char_p, P(tr_volume_root)]),
    "tr_volume_root_free": (None, [P(tr_volume_root)]),
    "tr_scene_add_hierarchical_volume": (_st, [_vp, P(tr_volume_root), P(tr_material), P(_u64)]),
    "tr_image_load": (_st, [ctypes.c_char_p, P(tr_image)]),
    "tr_image_free": (None, [P(tr_image)]),
    "tr_scene_bind_mesh_blob_textures": (_st, [_vp, P(tr_mesh_blob)]),
    "tr_load_obj_rt_geometry": (_st, [ctypes.c_char_p, _u32, P(tr_mesh_blob)]),
    "tr_mesh_blob_free": (None, [P(tr_mesh_blob)]),
    "tr_image_encode": (_st, [_u32, _u32, _u32, P(_u32), P(ctypes.c_uint8), _sz, P(_sz)]),
    "tr_image_save": (_st, [ctypes.c_char_p, _u32, _u32, P(_u32)]),
    "tr_scene_save_image": (_st, [_vp, ctypes.c_char_p]),
    "tr_scene_set_seed": (_st, [_vp, _u64]),
    "tr_scene_set_estimator": (_st, [_vp, _u32]),
    "tr_scene_set_frame_index": (_st, [_vp, _u32]),
    "tr_scene_trace_rays": (_st, [_vp, _sz, P(tr_ray), P(tr_hit)]),
    "tr_scene_ray_query": (_st, [_vp, P(ctypes.c_float), P(tr_sample_data), P(_u64)]),
    "tr_scene_trace_rays_noalloc": (_st, [_vp, _sz, _vp, _vp, _vp]),
    "tr_scene_render_paths": (_st, [_vp, _sz, P(_u32), P(ctypes.c_float), P(_u32)]),
    "tr_scene_get_stats": (_st, [_vp, P(tr_stats)]),
    "tr_scene_render_requests": (_st, [_vp, _sz, P(tr_request), P(tr_request_options), P(ctypes.c_float),
                                       P(ctypes.c_float), P(tr_request_stats)]),
    "tr_scene_set_radiance_cache": (_st, [_vp, P(tr_cache_settings)]),
    "tr_scene_get_radiance_cache": (_st, [_vp, P(_vp)]),
    "tr_scene_read_cache_records": (_st, [_vp, _sz, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float)]),
    "tr_scene_set_counting": (_st, [_vp, ctypes.c_int32]),
    "tr_scene_set_timing": (_st, [_vp, ctypes.c_int32]),
    "tr_mlp_create": (_st, [_u64, P(_vp)]),
    "tr_mlp_destroy": (None, [_vp]),
    "tr_mlp_get_params": (_st, [_vp, P(ctypes.c_float), _sz]),
    "tr_mlp_set_params": (_st, [_vp, P(ctypes.c_float), _sz]),
    "tr_mlp_infer": (_st, [_vp, _sz, P(ctypes.c_float), P(ctypes.c_float)]),
    "tr_mlp_gradient": (_st, [_vp, _sz, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                              P(ctypes.c_float)]),
    "tr_mlp_train_step": (_st, [_vp, _sz, P(ctypes.c_float), P(ctypes.c_float), P(ctypes.c_float),
                                P(ctypes.c_float)]),
    "tr_debug_build_scene_host": (_st, [_vp]),
    "tr_debug_trace_rays_host": (_st, [_vp, _sz, P(tr_ray), P(tr_hit), P(_u64), P(_u64)]),
    "tr_debug_bvh_check": (_st, [_vp, P(_u64)]),
}

_lib = None


class TempestError(RuntimeError):
    def __init__(self, status, message):
        super().__init__("tempest_rt status %d: %s" % (status, message))
        self.status = status


def load(path=None):
    """Loads libtempest_rt.so (built by `python -m tempestrenderer_amd.build`)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("libtempest_rt.so is not built (%s); run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status):
    if status != TR_OK:
        msg = load().tr_last_error_message()
        raise TempestError(status, msg.decode() if msg else "")
    return status

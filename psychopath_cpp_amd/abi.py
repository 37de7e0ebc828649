"""ctypes mirror of include/psy_gpu.h (structs and constants).

Shared by the GPU bindings (psychopath_cpp_amd.gpu) and the test oracle's
bindings, which take the same scene description.
"""
import ctypes as C

PSY_OK = 0
PSY_ERR_INVALID = 1
PSY_ERR_HIP = 2
PSY_ERR_UNSUPPORTED = 3
PSY_ERR_NO_DEVICE = 4

PSY_OBJ_SPHERE = 1
PSY_OBJ_BILINEAR = 2
PSY_OBJ_BICUBIC = 3
PSY_OBJ_SPHERE_LIGHT = 4
PSY_OBJ_RECTANGLE_LIGHT = 5
PSY_OBJ_SUBDIVISION = 6

PSY_INSTANCE_OBJECT = 0
PSY_INSTANCE_ASSEMBLY = 1

PSY_SHADER_LAMBERT = 1
PSY_SHADER_EMIT = 2
PSY_SHADER_GTR = 3
PSY_SCENE_RECT_LIGHT_FIX = 1
PSY_RENDER_SOBOL = 1

PSY_WRAY_PLANES = 20
PSY_RAY_PLANES = 15
PSY_RAY_OCCLUSION = 32
PSY_RAY_CAMERA = 1
PSY_RAY_R_DIFFUSE = 2
PSY_TRACE_OCCLUSION_ONLY = 1
PSY_TRACE_MODE_PLAIN = 0
PSY_TRACE_MODE_PERSISTENT = 1
PSY_TRACE_MODE_COHERENT = 2
PSY_TRACE_MODE_PLAIN_COHERENT = 3
PSY_RENDER_WAVEFRONT = 0
PSY_RENDER_FUSED = 1
PSY_RENDER_STREAM = 2

FLOATS_PER_KIND = {
    PSY_OBJ_SPHERE: 4,
    PSY_OBJ_BILINEAR: 12,
    PSY_OBJ_BICUBIC: 48,
    PSY_OBJ_SPHERE_LIGHT: 7,
    PSY_OBJ_RECTANGLE_LIGHT: 5,
}

MISS = 0xFFFFFFFF
PSY_STATIC_MAX_RAYS = 2097152
PSY_ISECT_PLANES = 44


class ObjectDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("time_samples", C.c_int32), ("data", C.POINTER(C.c_float)),
                ("n_vertices", C.c_int32), ("n_faces", C.c_int32),
                ("face_vert_counts", C.POINTER(C.c_int32)), ("face_vert_indices", C.POINTER(C.c_int32))]


class InstanceDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("data_index", C.c_int32), ("transform_index", C.c_int32),
                ("transform_count", C.c_int32), ("shader_index", C.c_int32)]


class ShaderDesc(C.Structure):
    _fields_ = [("kind", C.c_int32), ("color", C.c_float * 3), ("roughness", C.c_float), ("tail_shape", C.c_float),
                ("fresnel", C.c_float)]


class AssemblyDesc(C.Structure):
    _fields_ = [("objects", C.POINTER(ObjectDesc)), ("n_objects", C.c_int32),
                ("instances", C.POINTER(InstanceDesc)), ("n_instances", C.c_int32),
                ("transforms", C.POINTER(C.c_float)), ("n_transforms", C.c_int32),
                ("assemblies", C.POINTER(C.c_int32)), ("n_assemblies", C.c_int32)]


class CameraDesc(C.Structure):
    _fields_ = [("transforms", C.POINTER(C.c_float)), ("n_transforms", C.c_int32),
                ("fovs", C.POINTER(C.c_float)), ("n_fovs", C.c_int32),
                ("aperture_radii", C.POINTER(C.c_float)), ("n_aperture_radii", C.c_int32),
                ("focus_distances", C.POINTER(C.c_float)), ("n_focus_distances", C.c_int32)]


class RenderSettings(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("seed", C.c_int32),
                ("dice_rate", C.c_float), ("name", C.c_char * 256), ("output_path", C.c_char * 1024)]


class SceneDesc(C.Structure):
    _fields_ = [("assemblies", C.POINTER(AssemblyDesc)), ("n_assemblies", C.c_int32),
                ("shaders", C.POINTER(ShaderDesc)), ("n_shaders", C.c_int32),
                ("camera", CameraDesc), ("background", C.c_float * 3),
                ("dice_rate", C.c_float), ("min_upoly_size", C.c_float), ("flags", C.c_uint32)]


class Hits(C.Structure):
    _fields_ = [("t", C.c_void_p), ("inst", C.c_void_p), ("u", C.c_void_p), ("v", C.c_void_p),
                ("aux", C.c_void_p)]


class RenderParams(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("seed", C.c_uint32),
                ("path_length", C.c_int32), ("flags", C.c_uint32)]


class RenderTuning(C.Structure):
    _fields_ = [("mode", C.c_int32), ("trace_mode", C.c_int32), ("refill_min", C.c_int32), ("stack", C.c_int32),
                ("steps_per_graph", C.c_int32), ("trials", C.c_int32), ("pool", C.c_int64),
                ("dice_depth", C.c_int32), ("streams", C.c_int32),
                ("shade_waves", C.c_int32), ("mrays_per_s", C.c_double)]


class SceneInfo(C.Structure):
    _fields_ = [("n_leaf_instances", C.c_int64), ("n_objects", C.c_int64), ("n_bvh_nodes", C.c_int64),
                ("n_patches", C.c_int64), ("dice_depth", C.c_int32), ("bvh_stack_need", C.c_int32),
                ("device_bytes", C.c_int64), ("dice_cache_bytes", C.c_int64)]


# Every symbol the C ABI exports (include/psy_gpu.h).
EXPORTED_SYMBOLS = [
    "psy_ctx_create", "psy_ctx_destroy", "psy_ctx_set_trace_mode", "psy_ctx_retrace_count", "psy_ctx_trace_pass_counts", "psy_ctx_set_render_chunk", "psy_ctx_set_render_streams", "psy_ctx_set_render_shade_waves", "psy_ctx_set_render_mode", "psy_ctx_set_trace_timing", "psy_ctx_set_trace_stack", "psy_ctx_set_second_pass", "psy_ctx_set_continuation", "psy_ctx_set_stream_steps", "psy_render_autotune", "psy_render_adaptive", "psy_film_adaptive", "psy_bucket_size", "psy_ctx_trace_timing", "psy_last_error", "psy_abi_version",
    "psy_malloc", "psy_free", "psy_copy_to_device", "psy_copy_to_host",
    "psy_scene_create", "psy_scene_destroy", "psy_scene_get_info", "psy_scene_instance_ids", "psy_scene_set_dice_depth", "psy_subdivision_patches",
    "psy_trace_world", "psy_trace_rays", "psy_trace_rays_static", "psy_intersections", "psy_world_to_rays",
    "psy_parse_scene_count", "psy_parse_scene", "psy_parsed_scene_desc", "psy_parsed_scene_settings",
    "psy_parsed_scene_warnings", "psy_parsed_scene_destroy", "psy_parse_dump_tree",
    "psy_film_to_srgb8", "psy_write_png", "psy_write_pfm",
    "psy_halton", "psy_sobol_set_matrices", "psy_sobol", "psy_spectrum_set_grid", "psy_color_to_spectrum",
    "psy_pixel_samples", "psy_camera_rays", "psy_bounce_rays", "psy_render",
    "psy_scene_param_counts", "psy_scene_get_params", "psy_scene_set_params", "psy_adjoint",
]

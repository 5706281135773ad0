"""ctypes mirror of include/goodphotons.h (the C ABI of the HIP library).

Struct layouts here must match the header field for field; tests check the sizes
against the library's own sizeof() exports.
"""
from __future__ import annotations

import ctypes as C
import os

SPEC_CONST, SPEC_STEP, SPEC_GAUSS, SPEC_BLACKBODY, SPEC_SELLMEIER, SPEC_CAUCHY, SPEC_SAMPLED = range(7)
BVH_HOST_SAH, BVH_DEVICE_LBVH, BVH_PREBUILT = 0, 1, 2
ORDER_SHORTEST_FIRST, ORDER_ADMISSION = 0, 1
MAT_DIFFUSE, MAT_DIELECTRIC, MAT_MIRROR, MAT_GLOSSY = 0, 1, 2, 4

STATUS_NAMES = {
    0: "GP_OK", 1: "GP_ERR_INVALID", 2: "GP_ERR_UNSUPPORTED", 3: "GP_ERR_NO_DEVICE",
    4: "GP_ERR_HIP", 5: "GP_ERR_OOM", 6: "GP_ERR_CAPACITY", 7: "GP_ERR_REJECTED",
    8: "GP_ERR_STATE",
}


class Spectrum(C.Structure):
    _fields_ = [("kind", C.c_int32), ("table", C.c_int32), ("p", C.c_double * 6),
                ("band_lo", C.c_double), ("band_hi", C.c_double)]


class Material(C.Structure):
    _fields_ = [("type", C.c_int32), ("reserved", C.c_int32), ("roughness", C.c_double),
                ("reflect", Spectrum), ("ior", Spectrum)]


class Emitter(C.Structure):
    _fields_ = [("origin", C.c_double * 3), ("u", C.c_double * 3), ("v", C.c_double * 3),
                ("normal", C.c_double * 3), ("area", C.c_double), ("collimated", C.c_int32),
                ("reserved", C.c_int32), ("beam_dir", C.c_double * 3), ("spd", Spectrum)]


class BvhNode(C.Structure):
    _fields_ = [("lo", C.c_double * 3), ("hi", C.c_double * 3), ("left", C.c_int32), ("right", C.c_int32),
                ("first", C.c_int32), ("count", C.c_int32)]


class SceneDesc(C.Structure):
    _fields_ = [("n_tris", C.c_int64), ("n_materials", C.c_int32), ("n_emitters", C.c_int32),
                ("verts", C.POINTER(C.c_float)), ("mat_ids", C.POINTER(C.c_int32)),
                ("materials", C.POINTER(Material)), ("emitters", C.POINTER(Emitter)),
                ("spectrum_samples", C.POINTER(C.c_double)), ("n_spectrum_samples", C.c_int64),
                ("bvh_nodes", C.POINTER(BvhNode)), ("n_bvh_nodes", C.c_int64),
                ("bvh_prim_index", C.POINTER(C.c_int32)), ("n_bvh_prim_index", C.c_int64)]


class CameraDesc(C.Structure):
    _fields_ = [("eye", C.c_double * 3), ("target", C.c_double * 3), ("up", C.c_double * 3),
                ("fov_y_deg", C.c_double), ("res_x", C.c_int32), ("res_y", C.c_int32)]


class Energy(C.Structure):
    _fields_ = [("emitted", C.c_double), ("absorbed", C.c_double), ("sensor", C.c_double),
                ("escaped", C.c_double), ("residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TraceParams(C.Structure):
    _fields_ = [("n_photons", C.c_int64), ("seed_base", C.c_uint64), ("salt", C.c_uint64),
                ("hero_c", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("flags", C.c_int32)]


class TraceStats(C.Structure):
    _fields_ = [("photons", C.c_uint64), ("segments", C.c_uint64), ("connections", C.c_uint64),
                ("splats", C.c_uint64), ("iterations", C.c_uint64), ("wall_ms", C.c_double),
                ("kernel_ms", C.c_double * 6), ("kernel_launches", C.c_uint64 * 6),
                ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_tests", C.c_uint64),
                ("direct_connections", C.c_uint64)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["kernel_ms"] = list(self.kernel_ms)
        d["kernel_launches"] = list(self.kernel_launches)
        return d


class DeviceInfo(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
                ("wave_size", C.c_int32), ("total_mem", C.c_int64), ("free_mem", C.c_int64),
                ("lds_per_block", C.c_int32), ("reserved", C.c_int32)]


class SceneInfo(C.Structure):
    _fields_ = [("n_tris", C.c_int64), ("n_nodes", C.c_int64), ("bvh_depth", C.c_int32),
                ("n_emitters", C.c_int32), ("device_bytes", C.c_int64), ("build_ms", C.c_double),
                ("total_power", C.c_double)]


class Tuning(C.Structure):
    _fields_ = [("queue_capacity", C.c_int32), ("block_extend", C.c_int32),
                ("block_shade", C.c_int32), ("grid_waves_per_cu", C.c_int32),
                ("iters_per_graph", C.c_int32), ("use_graph", C.c_int32),
                ("lds_stack", C.c_int32), ("profile", C.c_int32),
                ("ray_sort", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("max_live_jobs", C.c_int32),
                ("order_policy", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", CameraDesc), ("params", TraceParams),
                ("user_tag", C.c_int64)]


class JobResult(C.Structure):
    _fields_ = [("job_id", C.c_int64), ("user_tag", C.c_int64), ("status", C.c_int32),
                ("res_x", C.c_int32), ("res_y", C.c_int32), ("energy", Energy),
                ("stats", TraceStats), ("submit_ms", C.c_double), ("start_ms", C.c_double),
                ("finish_ms", C.c_double)]


REFERENCE_SALT = 0x9E3779B97F4A7C15
TRACE_TRAVERSAL_STATS = 1           # gp_trace_params.flags: GP_TRACE_TRAVERSAL_STATS

# Every symbol include/goodphotons.h declares; the CPU tests check the built
# library exports each one.
EXPORTED_SYMBOLS = [
    "gp_abi_version", "gp_status_string", "gp_last_error", "gp_device_query",
    "gp_scene_create", "gp_scene_create_ex", "gp_scene_get_info", "gp_scene_destroy", "gp_mirror_planes",
    "gp_film_create", "gp_film_reset", "gp_film_download", "gp_film_download_raw", "gp_film_download_hits",
    "gp_film_seed",
    "gp_film_destroy", "gp_film_to_linear", "gp_film_tonemap", "gp_film_tonemap_noalloc",
    "gp_checkpoint_guard", "gp_film_save_checkpoint", "gp_film_load_checkpoint",
    "gp_trace", "gp_trace_shared", "gp_probe_closest_hit", "gp_probe_any_hit",
    "gp_get_tuning", "gp_set_tuning", "gp_autotune",
    "gp_server_create", "gp_server_submit", "gp_server_step", "gp_server_poll", "gp_server_take_result",
    "gp_server_pending", "gp_server_destroy",
]

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libgoodphotons.so")


def spectrum(kind, *params, band=None):
    s = Spectrum()
    s.kind = kind
    for i, v in enumerate(params):
        s.p[i] = float(v)
    if band is not None:
        s.band_lo, s.band_hi = float(band[0]), float(band[1])
    return s


def vec3(arr, vals):
    for i in range(3):
        arr[i] = float(vals[i])

KERNEL_NAMES = ["k_schedule", "k_emit", "k_extend", "k_shade", "k_connect", "k_bin"]

"""The C ABI library without a GPU: it loads, exports every symbol the header
declares, its ctypes mirror matches the C struct layouts, and calls that need a
device fail with a status instead of crashing."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

from goodphotons_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "goodphotons.h")


def test_header_declares_exactly_the_exported_list():
    text = open(HEADER).read()
    declared = set(re.findall(r"\b(gp_[a-z_]+)\s*\(", text))
    assert declared == set(abi.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    assert os.path.exists(abi.LIB_PATH), "build the HIP library first"
    out = subprocess.run(["nm", "-D", "--defined-only", abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (gp_[a-z_]+)$", out, re.M))
    missing = set(abi.EXPORTED_SYMBOLS) - exported
    assert not missing, missing
    lib = C.CDLL(abi.LIB_PATH)
    for name in abi.EXPORTED_SYMBOLS:
        assert getattr(lib, name) is not None


STRUCTS = ["gp_spectrum", "gp_material", "gp_emitter", "gp_bvh_node", "gp_scene_desc", "gp_camera_desc",
           "gp_energy", "gp_trace_params", "gp_trace_stats", "gp_device_info", "gp_scene_info",
           "gp_tuning", "gp_server_config", "gp_job_desc", "gp_job_result"]
CTYPES = {"gp_spectrum": abi.Spectrum, "gp_material": abi.Material, "gp_emitter": abi.Emitter,
          "gp_bvh_node": abi.BvhNode,
          "gp_scene_desc": abi.SceneDesc, "gp_camera_desc": abi.CameraDesc, "gp_energy": abi.Energy,
          "gp_trace_params": abi.TraceParams, "gp_trace_stats": abi.TraceStats,
          "gp_device_info": abi.DeviceInfo, "gp_scene_info": abi.SceneInfo, "gp_tuning": abi.Tuning,
          "gp_server_config": abi.ServerConfig, "gp_job_desc": abi.JobDesc,
          "gp_job_result": abi.JobResult}


def test_ctypes_layouts_match_c():
    src = "#include <stdio.h>\n#include <stddef.h>\n#include \"goodphotons.h\"\nint main(void){\n"
    for s in STRUCTS:
        src += f'printf("{s} %zu\\n", sizeof({s}));\n'
    src += 'printf("off_job_params %zu\\n", offsetof(gp_job_desc, params));\n'
    src += 'printf("off_result_stats %zu\\n", offsetof(gp_job_result, stats));\nreturn 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "sz.c")
        exe = os.path.join(d, "sz")
        open(c, "w").write(src)
        subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    sizes = dict(line.split() for line in out.strip().splitlines())
    for s in STRUCTS:
        assert int(sizes[s]) == C.sizeof(CTYPES[s]), s
    assert int(sizes["off_job_params"]) == abi.JobDesc.params.offset
    assert int(sizes["off_result_stats"]) == abi.JobResult.stats.offset


def test_no_device_is_a_status_not_a_crash():
    lib = C.CDLL(abi.LIB_PATH)
    lib.gp_device_query.argtypes = [C.POINTER(abi.DeviceInfo)]
    info = abi.DeviceInfo()
    st = lib.gp_device_query(C.byref(info))
    if st == 0:
        pytest.skip("a gfx950 device is present")
    assert st == 3
    lib.gp_last_error.restype = C.c_char_p
    assert lib.gp_last_error()
    lib.gp_film_create.argtypes = [C.c_int32, C.c_int32, C.c_int64, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    assert lib.gp_film_create(16, 16, 0, C.byref(h)) == 3


def test_invalid_arguments_rejected_before_device():
    lib = C.CDLL(abi.LIB_PATH)
    lib.gp_scene_create.argtypes = [C.c_void_p, C.POINTER(C.c_void_p)]
    assert lib.gp_scene_create(None, None) == 1
    lib.gp_set_tuning.argtypes = [C.POINTER(abi.Tuning)]
    t = abi.Tuning(queue_capacity=1000, grid_waves_per_cu=32, iters_per_graph=8, lds_stack=16)
    assert lib.gp_set_tuning(C.byref(t)) == 1
    t.queue_capacity = 1 << 16
    t.lds_stack = 20
    assert lib.gp_set_tuning(C.byref(t)) == 1


def test_python_host_fails_loudly_without_library(tmp_path):
    import goodphotons_amd as gp
    with pytest.raises(gp.GpError):
        gp.Library(str(tmp_path / "missing.so"))

"""The C ABI of the shared libraries, checked without a GPU.

- libtsp_hip.so exports every function include/tsp_hip.h declares; libtsp_host.so every function of
  include/tsp_scene.h (the scene builder).
- The C++ backend (SimBackend_HIP) reports a missing device the way the reference's backends
  report failures: an exception with the reason.
- The ctypes mirrors of the structs have the sizes and field offsets that a C compiler gives the
  headers (a probe compiled here with gcc against include/).
- Calls that reject their arguments before touching the device: a null context, null pointers.
"""
import ctypes
import os
import re
import subprocess

import pytest

from turbospharticles_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared(header):
    text = open(os.path.join(INCLUDE, header)).read()
    return sorted(set(re.findall(r"^[A-Za-z][\w \*]*?\b(tsp_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header,loader", [("tsp_hip.h", _capi.hip_lib), ("tsp_scene.h", _capi.host_lib)])
def test_every_declared_function_is_exported(header, loader):
    names = declared(header)
    assert names, header
    lib = loader()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "tsp_hip.h"
#include "tsp_scene.h"
#define S(T) printf(#T " size %zu\n", sizeof(T));
#define O(T, F) printf(#T " " #F " %zu\n", offsetof(T, F));
int main(void) {
  S(tsp_params) O(tsp_params, time_step) O(tsp_params, gravity) O(tsp_params, grid_cell_size_ratio)
  O(tsp_params, termination_cond)
  S(tsp_state_view) O(tsp_state_view, fluid) O(tsp_state_view, boundary) O(tsp_state_view, rigid_body)
  O(tsp_state_view, body) O(tsp_state_view, step) O(tsp_state_view, time)
  S(tsp_step_info) O(tsp_step_info, solver_iters) O(tsp_step_info, err_sum_raw) O(tsp_step_info, frame_dim)
  O(tsp_step_info, n_ff) O(tsp_step_info, list_bytes) O(tsp_step_info, rigid_force_torque)
  O(tsp_step_info, contact_iters) O(tsp_step_info, contacts)
  O(tsp_params, rigid_contact_density_dev_target) O(tsp_params, rigid_contact_max_solver_iters)
  S(tsp_export_view) O(tsp_export_view, cap_fluid) O(tsp_export_view, fluid_xyz) O(tsp_export_view, boundary_is_mesh)
  O(tsp_export_view, rigid_body)
  S(tsp_statistics) O(tsp_statistics, nbr_avg) O(tsp_statistics, nbr_p99) O(tsp_statistics, dens_rel_err_avg)
  O(tsp_statistics, vel_max)
  S(tsp_scene_spec) O(tsp_scene_spec, fluid_counts) O(tsp_scene_spec, jitter_seed) O(tsp_scene_spec, rigid_quat)
  O(tsp_scene_spec, rigid_omega)
  return 0;
}
"""


def test_struct_layouts_match_the_headers(tmp_path):
    src = tmp_path / "probe.c"
    exe = tmp_path / "probe"
    src.write_text(PROBE)
    subprocess.run(["gcc", "-std=c11", "-I", INCLUDE, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    mirrors = {"tsp_params": _capi.TspParams, "tsp_state_view": _capi.TspStateView,
               "tsp_step_info": _capi.TspStepInfo, "tsp_scene_spec": _capi.TspSceneSpec,
               "tsp_statistics": _capi.TspStatistics, "tsp_export_view": _capi.TspExportView}
    for line in out.strip().splitlines():
        parts = line.split()
        T, cls = parts[0], mirrors[parts[0]]
        if parts[1] == "size":
            assert ctypes.sizeof(cls) == int(parts[2]), line
        else:
            assert getattr(cls, parts[1]).offset == int(parts[2]), line


def test_calls_reject_a_null_context_without_a_device():
    lib = _capi.hip_lib()
    assert lib.tsp_set_params(None, None) == -1
    assert lib.tsp_step(None, 1) == -1
    assert lib.tsp_pull(None, None) == -1
    lib.tsp_last_error.restype = ctypes.c_char_p
    assert lib.tsp_last_error(None) == b"null context"
    lib.tsp_name.restype = ctypes.c_char_p
    assert lib.tsp_name() == b"HIP (gfx950)"
    lib.tsp_destroy(None)
    assert lib.tsp_create(0, None) == -1


def test_scene_builder_rejects_bad_specs():
    lib = _capi.host_lib()
    spec = _capi.TspSceneSpec()
    handle = ctypes.c_void_p()
    assert lib.tsp_scene_build(ctypes.byref(spec), ctypes.byref(handle)) != 0  # h = 0
    assert lib.tsp_scene_build(None, ctypes.byref(handle)) != 0


def test_scene_random_quat_is_unit_and_deterministic():
    lib = _capi.host_lib()
    q1 = (ctypes.c_float * 4)()
    q2 = (ctypes.c_float * 4)()
    lib.tsp_scene_random_quat(ctypes.c_uint64(7), q1)
    lib.tsp_scene_random_quat(ctypes.c_uint64(7), q2)
    assert list(q1) == list(q2)
    assert abs(sum(v * v for v in q1) - 1.0) < 1e-6


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present: the GPU tests run the backend instead")
def test_cpp_backend_reports_a_missing_device(tmp_path):
    exe = os.path.join(ROOT, "turbospharticles_amd", "lib", "sim_backend_roundtrip")
    r = subprocess.run([exe, str(tmp_path / "out.bin"), "4", "1"], capture_output=True, text=True)
    assert r.returncode == 3
    assert "no usable gfx950 device" in r.stderr


def test_batch_footprint_grows_with_every_population():
    lib = _capi.hip_lib()
    base = lib.tsp_batch_footprint(1000, 1000, 100, 1)
    assert base > 0
    for n in ((2000, 1000, 100, 1), (1000, 2000, 100, 1), (1000, 1000, 200, 1), (1000, 1000, 100, 2)):
        assert lib.tsp_batch_footprint(*n) > base, n
    # the fluid segment is rounded to the 256-particle block
    assert lib.tsp_batch_footprint(1, 0, 0, 0) == lib.tsp_batch_footprint(256, 0, 0, 0)


def test_batch_calls_reject_a_null_context():
    lib = _capi.hip_lib()
    assert lib.tsp_batch_step(None, 1) == -1
    assert lib.tsp_batch_remove_scene(None, 1) == -1
    assert lib.tsp_batch_update(None, 0, None, 0, None, None, None, None) == -1

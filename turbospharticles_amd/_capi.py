"""ctypes mirror of the C ABI in include/tsp_hip.h and include/tsp_scene.h.

The shared libraries are built in-tree by turbospharticles_amd.build:
  lib/libtsp_hip.so  -- HIP kernels for gfx950 and the tsp_* C ABI
  lib/libtsp_host.so -- host C++: scene builder and the SimBackend_HIP (I_SimBackend) adapter
Loading never falls back to anything: a missing library raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# TSP_HIP_LIB names another build of libtsp_hip.so (the variants the tuner measures)
HIP_LIB_PATH = os.environ.get("TSP_HIP_LIB") or os.path.join(LIB_DIR, "libtsp_hip.so")
HOST_LIB_PATH = os.path.join(LIB_DIR, "libtsp_host.so")

F_FIELDS = ["x", "y", "z", "vx", "vy", "vz", "vpx", "vpy", "vpz", "apx", "apy", "apz",
            "source", "diag", "pressure", "density"]
B_FIELDS = ["x", "y", "z", "vol", "psi"]
R_FIELDS = ["xrel", "yrel", "zrel", "vol", "psi"]
RB_FIELDS = ["xcm", "ycm", "zcm", "vx", "vy", "vz", "wx", "wy", "wz", "lx", "ly", "lz",
             "a00", "a01", "a02", "a10", "a11", "a12", "a20", "a21", "a22", "mass", "mass_inv",
             "i00", "i01", "i02", "i10", "i11", "i12", "i20", "i21", "i22"]

TSP_OK = 0
TSP_ERR_CAPACITY = -4
FLAG_MAX_ITERS = 1
FLAG_FRAME_REBUILT = 2
FLAG_RIGID_CONTACT = 4
FLAG_LIST_REGROWN = 8
FLAG_STEP_GRAPH = 16

_fp = ctypes.POINTER(ctypes.c_float)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


class TspParams(ctypes.Structure):
    _fields_ = [
        ("time_step", ctypes.c_float),
        ("particle_size", ctypes.c_float),
        ("kernel_radius", ctypes.c_float),
        ("kernel_radius_sq", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("mass", ctypes.c_float),
        ("viscosity_fluid", ctypes.c_float),
        ("viscosity_boundary", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("press_gamma", ctypes.c_float),
        ("jacobi_omega", ctypes.c_float),
        ("jacobi_warm_start", ctypes.c_float),
        ("density_dev_target", ctypes.c_float),
        ("grid_cell_size_ratio", ctypes.c_float),
        ("fixed_iters", ctypes.c_uint32),
        ("min_solver_iters", ctypes.c_uint32),
        ("max_solver_iters", ctypes.c_uint32),
        ("termination_cond", ctypes.c_uint32),
        ("rigid_contact_density_dev_target", ctypes.c_float),
        ("rigid_contact_max_solver_iters", ctypes.c_uint32),
    ]


class TspStateView(ctypes.Structure):
    _fields_ = [
        ("n_fluid", ctypes.c_uint32), ("n_boundary", ctypes.c_uint32),
        ("n_rigid", ctypes.c_uint32), ("n_bodies", ctypes.c_uint32),
        ("fluid_id", _u64p), ("fluid", _fp * len(F_FIELDS)),
        ("boundary_id", _u64p), ("boundary", _fp * len(B_FIELDS)),
        ("rigid_id", _u64p), ("rigid_body", _u32p), ("rigid", _fp * len(R_FIELDS)),
        ("body_id", _u32p), ("body_count", _u32p), ("body", _fp * len(RB_FIELDS)),
        ("step", ctypes.c_uint64), ("time", ctypes.c_double),
    ]


class TspStepInfo(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_uint64), ("time", ctypes.c_double),
        ("solver_iters", ctypes.c_uint32), ("flags", ctypes.c_uint32),
        ("avg_err_raw", ctypes.c_float), ("avg_err_pos", ctypes.c_float), ("avg_err_abs", ctypes.c_float),
        ("err_sum_raw", ctypes.c_double), ("err_sum_pos", ctypes.c_double), ("err_sum_abs", ctypes.c_double),
        ("sweeps_total", ctypes.c_uint64),
        ("frame_builds", ctypes.c_uint32),
        ("frame_orig", ctypes.c_int32 * 3), ("frame_dim", ctypes.c_uint32 * 3),
        ("n_ff", ctypes.c_uint64), ("n_fb", ctypes.c_uint64), ("n_fr", ctypes.c_uint64),
        ("n_rf", ctypes.c_uint64), ("n_rb", ctypes.c_uint64), ("n_rr", ctypes.c_uint64),
        ("list_bytes", ctypes.c_uint64),
        ("rigid_force_torque", ctypes.c_float * 6),
        ("contact_iters", ctypes.c_uint32), ("contact_avg_err", ctypes.c_float),
        ("contact_converged", ctypes.c_uint32), ("contact_boundary", ctypes.c_uint32),
        ("contacts", ctypes.c_uint64),
    ]


class TspExportView(ctypes.Structure):
    _fields_ = [
        ("step", ctypes.c_uint64), ("time", ctypes.c_double),
        ("cap_fluid", ctypes.c_uint32), ("n_fluid", ctypes.c_uint32),
        ("n_boundary", ctypes.c_uint32), ("n_rigid", ctypes.c_uint32),
        ("fluid_xyz", _fp), ("fluid_vel_mag", _fp),
        ("boundary_xyz", _fp), ("boundary_vol", _fp), ("boundary_is_mesh", ctypes.POINTER(ctypes.c_uint8)),
        ("rigid_xyz", _fp), ("rigid_vol", _fp), ("rigid_body", _u32p),
    ]


class TspMemory(ctypes.Structure):
    _fields_ = [("particles", ctypes.c_uint64 * 3), ("search", ctypes.c_uint64 * 3), ("lists", ctypes.c_uint64 * 3),
                ("misc", ctypes.c_uint64), ("alloc_bytes", ctypes.c_uint64), ("peak_alloc_bytes", ctypes.c_uint64)]


class TspStatistics(ctypes.Structure):
    _fields_ = [
        ("n_fluid", ctypes.c_uint32), ("n_boundary", ctypes.c_uint32), ("n_rigid", ctypes.c_uint32),
        ("solver_iters", ctypes.c_uint32),
        ("nbr_avg", ctypes.c_double * 6),
        ("nbr_min", ctypes.c_uint32 * 6), ("nbr_max", ctypes.c_uint32 * 6), ("nbr_p99", ctypes.c_uint32 * 6),
        ("pad", ctypes.c_uint32),
        ("dens_rel_err_avg", ctypes.c_double * 3),
        ("cfl", ctypes.c_double), ("vel_avg", ctypes.c_double), ("vel_min", ctypes.c_double),
        ("vel_max", ctypes.c_double),
    ]


class TspSceneSpec(ctypes.Structure):
    _fields_ = [
        ("particle_size", ctypes.c_float),
        ("rest_density", ctypes.c_float),
        ("boundary_layers", ctypes.c_uint32),
        ("container_origin", ctypes.c_float * 3),
        ("container_size", ctypes.c_float * 3),
        ("fluid_origin", ctypes.c_float * 3),
        ("fluid_counts", ctypes.c_uint32 * 3),
        ("jitter_frac", ctypes.c_float),
        ("jitter_seed", ctypes.c_uint64),
        ("rigid_enable", ctypes.c_uint32),
        ("rigid_center", ctypes.c_float * 3),
        ("rigid_size", ctypes.c_float * 3),
        ("rigid_density", ctypes.c_float),
        ("rigid_quat", ctypes.c_float * 4),
        ("rigid_vel", ctypes.c_float * 3),
        ("rigid_omega", ctypes.c_float * 3),
    ]


_hip = None
_host = None


def _declare_hip(lib):
    vp = ctypes.c_void_p
    i = ctypes.c_int
    sig = {
        "tsp_create": (i, [i, ctypes.POINTER(vp)]),
        "tsp_destroy": (None, [vp]),
        "tsp_last_error": (ctypes.c_char_p, [vp]),
        "tsp_name": (ctypes.c_char_p, []),
        "tsp_set_params": (i, [vp, ctypes.POINTER(TspParams)]),
        "tsp_set_domain": (i, [vp, _fp, _fp]),
        "tsp_set_cell_limit": (i, [vp, ctypes.c_uint64]),
        "tsp_set_step_graph": (i, [vp, i]),
        "tsp_pull": (i, [vp, ctypes.POINTER(TspStateView)]),
        "tsp_push": (i, [vp, ctypes.POINTER(TspStateView)]),
        "tsp_neighbor_search": (i, [vp]),
        "tsp_step": (i, [vp, ctypes.c_uint32]),
        "tsp_get_step_info": (i, [vp, ctypes.POINTER(TspStepInfo)]),
        "tsp_get_counts": (i, [vp, _u32p]),
        "tsp_get_neighbors": (i, [vp, i, _u32p, _u32p, ctypes.c_uint64, _u64p]),
        "tsp_get_sort": (i, [vp, _u32p, _u64p, _u32p, _u64p]),
        "tsp_get_rigid_world": (i, [vp, _fp, _fp]),
        "tsp_profile_enable": (i, [vp, i]),
        "tsp_profile_read": (i, [vp, ctypes.c_char_p, _u64p, ctypes.POINTER(ctypes.c_double), i]),
        "tsp_synchronize": (i, [vp]),
        "tsp_get_statistics": (i, [vp, ctypes.POINTER(TspStatistics)]),
        "tsp_get_statistics_noalloc": (i, [vp, ctypes.POINTER(TspStatistics)]),
        "tsp_get_memory": (i, [vp, ctypes.POINTER(TspMemory)]),
        "tsp_batch_scene_statistics": (i, [vp, ctypes.c_uint32, ctypes.POINTER(TspStatistics)]),
        "tsp_export_snapshot": (i, [vp, ctypes.c_uint32, ctypes.c_uint64, ctypes.POINTER(TspExportView)]),
        "tsp_batch_export_snapshot": (i, [vp, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint64,
                                          ctypes.POINTER(TspExportView)]),
        "tsp_batch_update": (i, [vp, ctypes.c_uint32, _u32p, ctypes.c_uint32, ctypes.POINTER(TspParams),
                                 ctypes.POINTER(ctypes.POINTER(TspStateView)), _fp, _u32p]),
        "tsp_batch_add_scene": (i, [vp, ctypes.POINTER(TspParams), ctypes.POINTER(TspStateView), _fp, _u32p]),
        "tsp_pull_spec": (i, [vp, ctypes.POINTER(TspSceneSpec)]),
        "tsp_batch_add_scene_spec": (i, [vp, ctypes.POINTER(TspParams), ctypes.POINTER(TspSceneSpec), _u32p]),
        "tsp_batch_update_specs": (i, [vp, ctypes.c_uint32, _u32p, ctypes.c_uint32, ctypes.POINTER(TspParams),
                                       ctypes.POINTER(ctypes.POINTER(TspSceneSpec)), _u32p]),
        "tsp_batch_remove_scene": (i, [vp, ctypes.c_uint32]),
        "tsp_batch_scene_ids": (i, [vp, _u32p, ctypes.c_uint32, _u32p]),
        "tsp_batch_scene_counts": (i, [vp, ctypes.c_uint32, _u32p]),
        "tsp_batch_push_scene": (i, [vp, ctypes.c_uint32, ctypes.POINTER(TspStateView)]),
        "tsp_batch_scene_info": (i, [vp, ctypes.c_uint32, ctypes.POINTER(TspStepInfo)]),
        "tsp_batch_step": (i, [vp, ctypes.c_uint32]),
        "tsp_batch_footprint": (ctypes.c_uint64, [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32,
                                                  ctypes.c_uint32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def _declare_host(lib):
    vp = ctypes.c_void_p
    i = ctypes.c_int
    sig = {
        "tsp_scene_build": (i, [ctypes.POINTER(TspSceneSpec), ctypes.POINTER(vp)]),
        "tsp_scene_counts": (None, [vp, _u32p]),
        "tsp_scene_domain": (None, [vp, _fp, _fp]),
        "tsp_scene_export": (i, [vp, ctypes.POINTER(TspStateView)]),
        "tsp_scene_free": (None, [vp]),
        "tsp_scene_random_quat": (None, [ctypes.c_uint64, _fp]),
        "tsp_scene_spec_counts": (i, [ctypes.POINTER(TspSceneSpec), _u32p]),
        "tsp_state_file_info": (i, [ctypes.c_char_p, _u32p, _u64p, ctypes.POINTER(ctypes.c_double)]),
        "tsp_state_file_read": (i, [ctypes.c_char_p, ctypes.POINTER(TspStateView)]),
        "tsp_state_file_write": (i, [ctypes.c_char_p, ctypes.POINTER(TspStateView)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip_lib():
    """libtsp_hip.so (HIP kernels + C ABI); raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB_PATH):
            raise RuntimeError("libtsp_hip.so is not built: run __graft_entry__.build() first")
        _hip = _declare_hip(ctypes.CDLL(HIP_LIB_PATH, mode=ctypes.RTLD_GLOBAL))
    return _hip


def host_lib():
    """libtsp_host.so (scene builder, SimBackend_HIP); raises if it has not been built."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB_PATH):
            raise RuntimeError("libtsp_host.so is not built: run __graft_entry__.build() first")
        _host = _declare_host(ctypes.CDLL(HOST_LIB_PATH, mode=ctypes.RTLD_GLOBAL))
    return _host


def fptr(a):
    return None if a is None else a.ctypes.data_as(_fp)


def u64ptr(a):
    return None if a is None else a.ctypes.data_as(_u64p)


def u32ptr(a):
    return None if a is None else a.ctypes.data_as(_u32p)


class HostState:
    """Host SoA arrays of one Sync::UniversalSimState (numpy, caller-owned)."""

    def __init__(self, n_fluid, n_boundary, n_rigid, n_bodies):
        self.n = (int(n_fluid), int(n_boundary), int(n_rigid), int(n_bodies))
        self.fluid_id = np.zeros(n_fluid, np.uint64)
        self.fluid = {k: np.zeros(n_fluid, np.float32) for k in F_FIELDS}
        self.boundary_id = np.zeros(n_boundary, np.uint64)
        self.boundary = {k: np.zeros(n_boundary, np.float32) for k in B_FIELDS}
        self.rigid_id = np.zeros(n_rigid, np.uint64)
        self.rigid_body = np.zeros(n_rigid, np.uint32)
        self.rigid = {k: np.zeros(n_rigid, np.float32) for k in R_FIELDS}
        self.body_id = np.zeros(n_bodies, np.uint32)
        self.body_count = np.zeros(n_bodies, np.uint32)
        self.body = {k: np.zeros(n_bodies, np.float32) for k in RB_FIELDS}
        self.step = 0
        self.time = 0.0
        self.domain = None

    def copy(self):
        c = HostState(*self.n)
        c.fluid_id[:] = self.fluid_id
        c.boundary_id[:] = self.boundary_id
        c.rigid_id[:] = self.rigid_id
        c.rigid_body[:] = self.rigid_body
        c.body_id[:] = self.body_id
        c.body_count[:] = self.body_count
        for src, dst in ((self.fluid, c.fluid), (self.boundary, c.boundary), (self.rigid, c.rigid),
                         (self.body, c.body)):
            for k in src:
                dst[k][:] = src[k]
        c.step, c.time = self.step, self.time
        c.domain = None if self.domain is None else (self.domain[0].copy(), self.domain[1].copy())
        return c

    def view(self):
        v = TspStateView()
        v.n_fluid, v.n_boundary, v.n_rigid, v.n_bodies = self.n
        v.fluid_id = u64ptr(self.fluid_id)
        for i, k in enumerate(F_FIELDS):
            v.fluid[i] = fptr(self.fluid[k])
        v.boundary_id = u64ptr(self.boundary_id)
        for i, k in enumerate(B_FIELDS):
            v.boundary[i] = fptr(self.boundary[k])
        v.rigid_id = u64ptr(self.rigid_id)
        v.rigid_body = u32ptr(self.rigid_body)
        for i, k in enumerate(R_FIELDS):
            v.rigid[i] = fptr(self.rigid[k])
        v.body_id = u32ptr(self.body_id)
        v.body_count = u32ptr(self.body_count)
        for i, k in enumerate(RB_FIELDS):
            v.body[i] = fptr(self.body[k])
        v.step = self.step
        v.time = self.time
        return v


def save_state(path, state):
    """Writes a HostState as the reference's TSPSIM2 simulation state file."""
    v = state.view()
    rc = host_lib().tsp_state_file_write(os.fsencode(path), ctypes.byref(v))
    if rc != TSP_OK:
        raise RuntimeError(f"tsp_state_file_write failed ({rc}): {path}")


def load_state(path):
    """Reads a TSPSIM2 simulation state file into a new HostState (no domain)."""
    lib = host_lib()
    counts = (ctypes.c_uint32 * 4)()
    rc = lib.tsp_state_file_info(os.fsencode(path), counts, None, None)
    if rc != TSP_OK:
        raise RuntimeError(f"tsp_state_file_info failed ({rc}): {path}")
    st = HostState(*counts)
    v = st.view()
    rc = lib.tsp_state_file_read(os.fsencode(path), ctypes.byref(v))
    if rc != TSP_OK:
        raise RuntimeError(f"tsp_state_file_read failed ({rc}): {path}")
    st.step, st.time = int(v.step), float(v.time)
    return st


class ExportSnapshot:
    """Sync::ExportSnapshot as numpy arrays (fluid_xyz is n x 3 in the file's triple layout)."""

    def __init__(self, step=0, time=0.0, fluid_xyz=None, fluid_vel_mag=None, boundary_xyz=None,
                 boundary_vol=None, boundary_is_mesh=None, rigid_xyz=None, rigid_vol=None, rigid_body=None):
        f32 = np.float32
        self.step, self.time = int(step), float(time)
        self.fluid_xyz = np.zeros((0, 3), f32) if fluid_xyz is None else np.ascontiguousarray(fluid_xyz, f32)
        self.fluid_vel_mag = np.zeros(0, f32) if fluid_vel_mag is None else np.ascontiguousarray(fluid_vel_mag, f32)
        self.boundary_xyz = None if boundary_xyz is None else np.ascontiguousarray(boundary_xyz, f32)
        self.boundary_vol = None if boundary_vol is None else np.ascontiguousarray(boundary_vol, f32)
        self.boundary_is_mesh = None if boundary_is_mesh is None else np.ascontiguousarray(boundary_is_mesh, np.uint8)
        self.rigid_xyz = np.zeros((0, 3), f32) if rigid_xyz is None else np.ascontiguousarray(rigid_xyz, f32)
        self.rigid_vol = np.zeros(0, f32) if rigid_vol is None else np.ascontiguousarray(rigid_vol, f32)
        self.rigid_body = np.zeros(0, np.uint32) if rigid_body is None else np.ascontiguousarray(rigid_body, np.uint32)

    def view(self):
        v = TspExportView()
        v.step, v.time = self.step, self.time
        v.n_fluid = v.cap_fluid = self.fluid_vel_mag.size
        v.n_rigid = self.rigid_vol.size
        v.fluid_xyz, v.fluid_vel_mag = fptr(self.fluid_xyz), fptr(self.fluid_vel_mag)
        v.rigid_xyz, v.rigid_vol, v.rigid_body = fptr(self.rigid_xyz), fptr(self.rigid_vol), u32ptr(self.rigid_body)
        if self.boundary_xyz is not None:
            v.n_boundary = self.boundary_xyz.shape[0]
            v.boundary_xyz, v.boundary_vol = fptr(self.boundary_xyz), fptr(self.boundary_vol)
            if self.boundary_is_mesh is not None:
                v.boundary_is_mesh = self.boundary_is_mesh.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        return v


def export_from(call, n_fluid_cap, n_boundary, n_rigid, include_boundary):
    """Runs an export call (call(view) -> status) into new arrays, growing the fluid arrays once
    if the surface set is larger than n_fluid_cap."""
    cap = max(int(n_fluid_cap), 1)
    for _ in range(2):
        snap = ExportSnapshot(fluid_xyz=np.zeros((cap, 3), np.float32), fluid_vel_mag=np.zeros(cap, np.float32),
                              rigid_xyz=np.zeros((n_rigid, 3), np.float32), rigid_vol=np.zeros(n_rigid, np.float32),
                              rigid_body=np.zeros(n_rigid, np.uint32))
        if include_boundary:
            snap.boundary_xyz = np.zeros((n_boundary, 3), np.float32)
            snap.boundary_vol = np.zeros(n_boundary, np.float32)
            snap.boundary_is_mesh = np.zeros(n_boundary, np.uint8)
        v = snap.view()
        v.cap_fluid = cap
        if not include_boundary:
            v.boundary_xyz = None
        rc = call(v)
        if rc == TSP_ERR_CAPACITY and v.n_fluid > cap:
            cap = v.n_fluid
            continue
        if rc != TSP_OK:
            return rc, None
        snap.step, snap.time = int(v.step), float(v.time)
        snap.fluid_xyz = snap.fluid_xyz[:v.n_fluid]
        snap.fluid_vel_mag = snap.fluid_vel_mag[:v.n_fluid]
        return rc, snap
    return TSP_ERR_CAPACITY, None


